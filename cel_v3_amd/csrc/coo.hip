// coo.hip -- COO ordering and COO -> CSR in HBM.
//
// cel_gpu_coo_sort replaces cel_sp_mat_format_sparse_31_sort
// (src/fspmt/cel_sp_mat_module.f90:518-585); cel_gpu_coo_row_offsets and
// cel_gpu_mat_create_coo_sorted replace cel_sp_mat_to_csr (:430-456) and the
// matrix build after it.
//
// The reference sorts on one CPU thread: a stable merge sort by row, then a
// loop that walks the row-sorted entries and merge-sorts each segment it
// closes by column.  Both stages are stable sorts of contiguous segments,
// so on the GPU each becomes one stable LSD radix sort (rocPRIM through
// hipCUB) over packed keys:
//   CEL_SORT_ROW_COL    key = row << cb | col                 (one sort)
//   CEL_SORT_CEL_COMPAT key = row                              (sort 1)
//                       key = segment << cb | col, the input   (sort 2)
//                       in the order sort 1 left it
// where `segment` numbers the contiguous groups that the reference's loop
// column-sorts; every entry after the last group it reaches is a group of
// its own, so it keeps its place.  The groups follow from the row-sorted
// rows alone (transitions, below), computed with two scans.  Radix sort is
// stable, and a stable sort's output is fixed by its keys and its input
// order, so the result is the reference's, entry for entry.
//
// The reference's loop, restated on 0-based positions p of the row-sorted
// entries (row r[p], 0-based local; R = r + row_offset + 1 is the value the
// reference holds):
//   p is a transition when p == 0 and (R[0] != 1 or nnz == 1), or when
//   p > 0 and (r[p] != r[p-1] or p == nnz-1);
//   the loop handles the first K = nrows-1 transitions and stops;
//   transition p closes the segment ending at e(p) = nnz-1 if p == nnz-1,
//   else max(p-1, 0); segments are consecutive, so entry q belongs to the
//   group numbered by how many closing ends lie before q.
// With fewer than K transitions the reference reads rows(nnz+1), past its
// array: CEL_ERR_REF_UNDEFINED, and the entries are left row-sorted.
#include <hipcub/hipcub.hpp>

#include "cel_internal.hpp"

namespace cel {
namespace {

constexpr int kThreads = 256;

int grid_for(long long n)
{
    long long g = (n + kThreads - 1) / kThreads;
    long long cap = (long long)rt().num_cus * 16;
    return (int)(g < 1 ? 1 : g > cap ? cap : g);
}

int bits_for(unsigned long long v)  // bits to hold 0..v
{
    int b = 0;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
}

// Local 0-based row/col of every entry, range check, packed keys, identity.
__global__ void __launch_bounds__(kThreads)
k_coo_keys(long long nnz, const int* __restrict__ rows, const int* __restrict__ cols, int rbase,
           int nrows, int cbase, int ncols, int cb, int mode, uint64_t* __restrict__ key,
           uint32_t* __restrict__ rkey, int* __restrict__ idx, int* __restrict__ bad)
{
    long long stride = (long long)gridDim.x * blockDim.x;
    int any = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += stride) {
        long long r = (long long)rows[i] - rbase;
        long long c = (long long)cols[i] - cbase;
        bool ok = r >= 0 && r < nrows && c >= 0 && c < ncols;
        any |= !ok;
        uint32_t ru = ok ? (uint32_t)r : 0u, cu = ok ? (uint32_t)c : 0u;
        if (mode == CEL_SORT_ROW_COL)
            key[i] = ((uint64_t)ru << cb) | cu;
        else
            rkey[i] = ru;
        idx[i] = (int)i;
    }
    if (any) atomicOr(bad, 1);
}

// t[p]: is p a transition of the reference's column loop.
__global__ void __launch_bounds__(kThreads)
k_transitions(long long nnz, const uint32_t* __restrict__ r, long long row_offset,
              int* __restrict__ t)
{
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += stride) {
        int v;
        if (p == 0)
            v = nnz == 1 || ((long long)r[0] + row_offset) != 0;
        else
            v = r[p] != r[p - 1] || p == nnz - 1;
        t[p] = v;
    }
}

// The position of the K-th transition (ord is the inclusive count).
__global__ void __launch_bounds__(kThreads)
k_kth(long long nnz, const int* __restrict__ t, const int* __restrict__ ord, int K,
      int* __restrict__ pk)
{
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += stride)
        if (t[p] && ord[p] == K) *pk = (int)p;
}

// mark[q] = 1 when q closes a handled segment or lies after the last one.
__global__ void __launch_bounds__(kThreads)
k_group_ends(long long nnz, const int* __restrict__ t, const int* __restrict__ ord, int K,
             const int* __restrict__ pk, int* __restrict__ mark)
{
    long long pK = *pk;
    long long endK = pK == nnz - 1 ? nnz - 1 : (pK > 1 ? pK - 1 : 0);
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nnz; q += stride) {
        bool handled_next = q + 1 <= nnz - 2 && t[q + 1] && ord[q + 1] <= K;  // e(q+1) = q
        bool last = q == nnz - 1 && t[q] && ord[q] <= K;                     // e = nnz-1
        bool first = q == 0 && t[0] && ord[0] <= K;                          // e(0) = 0
        mark[q] = (handled_next || last || first || q > endK) ? 1 : 0;
    }
}

// key = group << cb | col of the entry now at p (group = exclusive count).
__global__ void __launch_bounds__(kThreads)
k_group_keys(long long nnz, const int* __restrict__ seg, const int* __restrict__ perm,
             const int* __restrict__ cols, int cbase, int cb, uint64_t* __restrict__ key)
{
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += stride)
        key[p] = ((uint64_t)(uint32_t)seg[p] << cb) | (uint32_t)(cols[perm[p]] - cbase);
}

__global__ void __launch_bounds__(kThreads)
k_gather3(long long nnz, const int* __restrict__ perm, const int* __restrict__ rin,
          const int* __restrict__ cin, const double* __restrict__ vin, int* __restrict__ rout,
          int* __restrict__ cout, double* __restrict__ vout)
{
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += stride) {
        int s = perm[p];
        rout[p] = rin[s];
        cout[p] = cin[s];
        vout[p] = vin[s];
    }
}

// Rows of consecutive entries must not decrease for a CSR build.
__global__ void __launch_bounds__(kThreads)
k_rows_monotone(long long nnz, const int* __restrict__ rows, int* __restrict__ bad)
{
    long long stride = (long long)gridDim.x * blockDim.x;
    int any = 0;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x + 1; p < nnz; p += stride)
        any |= rows[p] < rows[p - 1];
    if (any) atomicOr(bad, 1);
}

// cel_sp_mat_to_csr: counts per row, then a scan.  cnt has nrows+1 slots,
// cnt[0] = 0; entries outside [0, nrows) are counted nowhere (the caller
// has range-checked them).
__global__ void __launch_bounds__(kThreads)
k_row_count(long long nnz, const int* __restrict__ rows, int rbase, int nrows,
            int* __restrict__ cnt)
{
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += stride) {
        long long r = (long long)rows[p] - rbase;
        if (r >= 0 && r < nrows) atomicAdd(&cnt[r + 1], 1);
    }
}

__global__ void __launch_bounds__(kThreads)
k_rows_in_range(long long nnz, const int* __restrict__ rows, int rbase, int nrows,
                int* __restrict__ bad)
{
    long long stride = (long long)gridDim.x * blockDim.x;
    int any = 0;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += stride) {
        long long r = (long long)rows[p] - rbase;
        any |= r < 0 || r >= nrows;
    }
    if (any) atomicOr(bad, 1);
}

__global__ void __launch_bounds__(kThreads)
k_fill_int(long long n, int* __restrict__ a, int v)
{
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) a[i] = v;
}

__global__ void __launch_bounds__(kThreads)
k_sub_int(long long n, const int* a, int v, int* out)  // may run in place
{
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = a[i] - v;
}

struct SortBufs {
    DevBuf<int> rin, cin, idx, perm, a, b, c, flag;
    DevBuf<double> vin;
    DevBuf<uint64_t> k0, k1;
    DevBuf<unsigned char> temp;
};

int temp_fit(SortBufs& S, size_t bytes)
{
    if (S.temp.n >= bytes && S.temp.p) return CEL_OK;
    CEL_HIP(hipStreamSynchronize(rt().stream));  // earlier passes may still read it
    return S.temp.alloc(bytes ? bytes : 1);
}

// Sorts (rows, cols, vals) on the device into (ro, co, vo); inputs are
// device copies owned by S (rin/cin/vin).  Returns CEL_OK or
// CEL_ERR_REF_UNDEFINED (outputs then row-sorted).
int sort_device(SortBufs& S, int nrows, int ncols, long long nnz, int rbase, int cbase,
                long long row_offset, int mode, int* ro, int* co, double* vo)
{
    hipStream_t st = rt().stream;
    const int g = grid_for(nnz);
    const int n = (int)nnz;
    const int cb = bits_for(ncols > 0 ? (unsigned long long)(ncols - 1) : 0ULL);
    const int rb = bits_for(nrows > 0 ? (unsigned long long)(nrows - 1) : 0ULL);
    CEL_TRY(S.idx.alloc(nnz));
    CEL_TRY(S.perm.alloc(nnz));
    CEL_TRY(S.k0.alloc(nnz));
    CEL_TRY(S.k1.alloc(nnz));
    CEL_TRY(S.flag.alloc(2));
    CEL_HIP(hipMemsetAsync(S.flag.p, 0, 2 * sizeof(int), st));
    uint32_t* rk0 = reinterpret_cast<uint32_t*>(S.k0.p);
    uint32_t* rk1 = reinterpret_cast<uint32_t*>(S.k1.p);
    k_coo_keys<<<g, kThreads, 0, st>>>(nnz, S.rin.p, S.cin.p, rbase, nrows, cbase, ncols, cb, mode,
                                        S.k0.p, rk0, S.idx.p, S.flag.p);
    CEL_HIP(hipGetLastError());
    int bad = 0;
    CEL_HIP(hipMemcpyAsync(&bad, S.flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    CEL_HIP(hipStreamSynchronize(st));
    if (bad) return fail(CEL_ERR_BAD_MATRIX, "row or column index out of range");

    size_t bytes = 0;
    if (mode == CEL_SORT_ROW_COL) {
        const int eb = cb + rb > 0 ? cb + rb : 1;
        CEL_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, S.k0.p, S.k1.p, S.idx.p, S.perm.p,
                                                   n, 0, eb, st));
        CEL_TRY(temp_fit(S, bytes));
        CEL_HIP(hipcub::DeviceRadixSort::SortPairs(S.temp.p, bytes, S.k0.p, S.k1.p, S.idx.p, S.perm.p,
                                                   n, 0, eb, st));
        k_gather3<<<g, kThreads, 0, st>>>(nnz, S.perm.p, S.rin.p, S.cin.p, S.vin.p, ro, co, vo);
        CEL_HIP(hipGetLastError());
        return CEL_OK;
    }

    // stage 1: stable by row
    CEL_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, rk0, rk1, S.idx.p, S.perm.p, n, 0,
                                               rb > 0 ? rb : 1, st));
    CEL_TRY(temp_fit(S, bytes));
    CEL_HIP(hipcub::DeviceRadixSort::SortPairs(S.temp.p, bytes, rk0, rk1, S.idx.p, S.perm.p, n, 0,
                                               rb > 0 ? rb : 1, st));
    const int K = nrows - 1;
    if (K <= 0) {  // the loop body never runs
        k_gather3<<<g, kThreads, 0, st>>>(nnz, S.perm.p, S.rin.p, S.cin.p, S.vin.p, ro, co, vo);
        CEL_HIP(hipGetLastError());
        return CEL_OK;
    }
    CEL_TRY(S.a.alloc(nnz));
    CEL_TRY(S.b.alloc(nnz));
    CEL_TRY(S.c.alloc(nnz));
    k_transitions<<<g, kThreads, 0, st>>>(nnz, rk1, row_offset, S.a.p);
    CEL_HIP(hipGetLastError());
    CEL_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, S.a.p, S.b.p, n, st));
    CEL_TRY(temp_fit(S, bytes));
    CEL_HIP(hipcub::DeviceScan::InclusiveSum(S.temp.p, bytes, S.a.p, S.b.p, n, st));
    int total = 0;
    CEL_HIP(hipMemcpyAsync(&total, S.b.p + (nnz - 1), sizeof(int), hipMemcpyDeviceToHost, st));
    CEL_HIP(hipStreamSynchronize(st));
    if (total < K) {
        k_gather3<<<g, kThreads, 0, st>>>(nnz, S.perm.p, S.rin.p, S.cin.p, S.vin.p, ro, co, vo);
        CEL_HIP(hipGetLastError());
        CEL_HIP(hipStreamSynchronize(st));
        return fail(CEL_ERR_REF_UNDEFINED,
                    "the reference's column loop needs nrows-1 row transitions and finds " +
                        std::to_string(total) + ": it reads past its arrays (empty rows)");
    }
    k_kth<<<g, kThreads, 0, st>>>(nnz, S.a.p, S.b.p, K, S.flag.p + 1);
    k_group_ends<<<g, kThreads, 0, st>>>(nnz, S.a.p, S.b.p, K, S.flag.p + 1, S.c.p);
    CEL_HIP(hipGetLastError());
    CEL_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, S.c.p, S.a.p, n, st));
    CEL_TRY(temp_fit(S, bytes));
    CEL_HIP(hipcub::DeviceScan::ExclusiveSum(S.temp.p, bytes, S.c.p, S.a.p, n, st));
    // stage 2: stable by (group, col) over the row-sorted order
    k_group_keys<<<g, kThreads, 0, st>>>(nnz, S.a.p, S.perm.p, S.cin.p, cbase, cb, S.k0.p);
    CEL_HIP(hipGetLastError());
    const int sb = bits_for((unsigned long long)nnz);
    CEL_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, S.k0.p, S.k1.p, S.perm.p, S.idx.p,
                                               n, 0, cb + sb, st));
    CEL_TRY(temp_fit(S, bytes));
    CEL_HIP(hipcub::DeviceRadixSort::SortPairs(S.temp.p, bytes, S.k0.p, S.k1.p, S.perm.p, S.idx.p,
                                               n, 0, cb + sb, st));
    k_gather3<<<g, kThreads, 0, st>>>(nnz, S.idx.p, S.rin.p, S.cin.p, S.vin.p, ro, co, vo);
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

int copy_in(SortBufs& S, long long nnz, const int* rows, const int* cols, const double* vals,
            int memspace)
{
    CEL_TRY(S.rin.alloc(nnz));
    CEL_TRY(S.cin.alloc(nnz));
    CEL_TRY(S.vin.alloc(nnz));
    hipMemcpyKind k = memspace == CEL_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    hipStream_t st = rt().stream;
    CEL_HIP(hipMemcpyAsync(S.rin.p, rows, sizeof(int) * nnz, k, st));
    CEL_HIP(hipMemcpyAsync(S.cin.p, cols, sizeof(int) * nnz, k, st));
    CEL_HIP(hipMemcpyAsync(S.vin.p, vals, sizeof(double) * nnz, k, st));
    return CEL_OK;
}

bool bad_common(int nrows, int ncols, int nnz, int index_base, int mode, int memspace)
{
    return nrows < 0 || ncols < 0 || nnz < 0 || (index_base != 0 && index_base != 1) ||
           (mode != CEL_SORT_ROW_COL && mode != CEL_SORT_CEL_COMPAT) ||
           (memspace != CEL_MEM_HOST && memspace != CEL_MEM_DEVICE);
}

}  // namespace

// cel_sp_mat_to_csr on the device: xadj[0] = out_base, xadj[i+1] - xadj[i]
// = entries of row i.  Rows outside [0, nrows) are an error (the reference
// writes past its count array there).
int row_offsets_device(int nrows, long long nnz, const int* rows, int rbase, int out_base, int* xadj)
{
    hipStream_t st = rt().stream;
    DevBuf<int> flag, cnt;
    CEL_TRY(flag.alloc(1));
    CEL_TRY(cnt.alloc((size_t)nrows + 1));
    CEL_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), st));
    k_fill_int<<<grid_for(nrows + 1LL), kThreads, 0, st>>>(nrows + 1LL, cnt.p, 0);
    k_fill_int<<<1, 1, 0, st>>>(1, cnt.p, out_base);  // xadj[0] = out_base carries through the scan
    if (nnz > 0) {
        k_rows_in_range<<<grid_for(nnz), kThreads, 0, st>>>(nnz, rows, rbase, nrows, flag.p);
        k_row_count<<<grid_for(nnz), kThreads, 0, st>>>(nnz, rows, rbase, nrows, cnt.p);
    }
    CEL_HIP(hipGetLastError());
    int bad = 0;
    CEL_HIP(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    size_t bytes = 0;
    CEL_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, cnt.p, xadj, nrows + 1, st));
    DevBuf<unsigned char> temp;
    CEL_TRY(temp.alloc(bytes ? bytes : 1));
    CEL_HIP(hipcub::DeviceScan::InclusiveSum(temp.p, bytes, cnt.p, xadj, nrows + 1, st));
    CEL_HIP(hipStreamSynchronize(st));
    if (bad) return fail(CEL_ERR_BAD_MATRIX, "row index out of range");
    return CEL_OK;
}

int launch_coo_to_csr(int nrows, int nnz, const int* rows0, int* xadj)
{
    return row_offsets_device(nrows, nnz, rows0, 0, 0, xadj);
}

}  // namespace cel

using namespace cel;

extern "C" int cel_gpu_coo_sort(int nrows, int ncols, int nnz, int* rows, int* cols, double* vals,
                                int index_base, int row_offset, int col_offset, int mode,
                                int memspace)
{
    CEL_TRY(ensure_init());
    if (bad_common(nrows, ncols, nnz, index_base, mode, memspace) ||
        (nnz > 0 && (!rows || !cols || !vals)))
        return fail(CEL_ERR_ARG, "bad arguments to cel_gpu_coo_sort");
    if (nnz == 0) {
        if (mode == CEL_SORT_CEL_COMPAT && nrows > 1)
            return fail(CEL_ERR_REF_UNDEFINED, "no entries: the reference's column loop reads rows(1)");
        return CEL_OK;
    }
    SortBufs S;
    CEL_TRY(copy_in(S, nnz, rows, cols, vals, memspace));
    const int rbase = index_base + row_offset, cbase = index_base + col_offset;
    int rc;
    if (memspace == CEL_MEM_DEVICE) {
        rc = sort_device(S, nrows, ncols, nnz, rbase, cbase, row_offset, mode, rows, cols, vals);
    } else {
        DevBuf<int> ro, co;
        DevBuf<double> vo;
        CEL_TRY(ro.alloc(nnz));
        CEL_TRY(co.alloc(nnz));
        CEL_TRY(vo.alloc(nnz));
        rc = sort_device(S, nrows, ncols, nnz, rbase, cbase, row_offset, mode, ro.p, co.p, vo.p);
        if (rc == CEL_OK || rc == CEL_ERR_REF_UNDEFINED) {
            hipStream_t st = rt().stream;
            CEL_HIP(hipMemcpyAsync(rows, ro.p, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToHost, st));
            CEL_HIP(hipMemcpyAsync(cols, co.p, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToHost, st));
            CEL_HIP(hipMemcpyAsync(vals, vo.p, sizeof(double) * (size_t)nnz, hipMemcpyDeviceToHost, st));
        }
    }
    CEL_HIP(hipStreamSynchronize(rt().stream));
    return rc;
}

extern "C" int cel_gpu_coo_row_offsets(int nrows, int nnz, const int* rows, int index_base,
                                       int row_offset, int* xadj, int memspace)
{
    CEL_TRY(ensure_init());
    if (nrows < 0 || nnz < 0 || (index_base != 0 && index_base != 1) || !xadj ||
        (nnz > 0 && !rows) || (memspace != CEL_MEM_HOST && memspace != CEL_MEM_DEVICE))
        return fail(CEL_ERR_ARG, "bad arguments to cel_gpu_coo_row_offsets");
    hipStream_t st = rt().stream;
    DevBuf<int> rin, dx;
    const int* r = rows;
    if (memspace == CEL_MEM_HOST && nnz > 0) {
        CEL_TRY(rin.alloc(nnz));
        CEL_HIP(hipMemcpyAsync(rin.p, rows, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice, st));
        r = rin.p;
    }
    int* x = xadj;
    if (memspace == CEL_MEM_HOST) {
        CEL_TRY(dx.alloc((size_t)nrows + 1));
        x = dx.p;
    }
    CEL_TRY(row_offsets_device(nrows, nnz, r, index_base + row_offset, index_base, x));
    if (memspace == CEL_MEM_HOST)
        CEL_HIP(hipMemcpyAsync(xadj, x, sizeof(int) * ((size_t)nrows + 1), hipMemcpyDeviceToHost, st));
    CEL_HIP(hipStreamSynchronize(st));
    return CEL_OK;
}

extern "C" int cel_gpu_mat_create_coo_sorted(cel_gpu_mat** A, int nrows, int ncols, int nnz,
                                             const int* rows, const int* cols, const double* vals,
                                             int index_base, int row_offset, int col_offset,
                                             int mode, int memspace)
{
    CEL_TRY(ensure_init());
    if (!A || bad_common(nrows, ncols, nnz, index_base, mode, memspace) ||
        (nnz > 0 && (!rows || !cols || !vals)))
        return fail(CEL_ERR_ARG, "bad arguments to cel_gpu_mat_create_coo_sorted");
    if (nnz == 0 && mode == CEL_SORT_CEL_COMPAT && nrows > 1)
        return fail(CEL_ERR_REF_UNDEFINED, "no entries: the reference's column loop reads rows(1)");
    hipStream_t st = rt().stream;
    const int rbase = index_base + row_offset, cbase = index_base + col_offset;
    SortBufs S;
    DevBuf<int> ro, co, xadj;
    DevBuf<double> vo;
    CEL_TRY(ro.alloc(nnz > 0 ? nnz : 1));
    CEL_TRY(co.alloc(nnz > 0 ? nnz : 1));
    CEL_TRY(vo.alloc(nnz > 0 ? nnz : 1));
    CEL_TRY(xadj.alloc((size_t)nrows + 1));
    if (nnz > 0) {
        CEL_TRY(copy_in(S, nnz, rows, cols, vals, memspace));
        CEL_TRY(sort_device(S, nrows, ncols, nnz, rbase, cbase, row_offset, mode, ro.p, co.p, vo.p));
        // a CSR build needs rows that never decrease; the reference's sort
        // can leave a last row of one entry inside the row before it
        CEL_HIP(hipMemsetAsync(S.flag.p, 0, sizeof(int), st));
        k_rows_monotone<<<grid_for(nnz), kThreads, 0, st>>>(nnz, ro.p, S.flag.p);
        CEL_HIP(hipGetLastError());
        int bad = 0;
        CEL_HIP(hipMemcpyAsync(&bad, S.flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
        CEL_HIP(hipStreamSynchronize(st));
        if (bad)
            return fail(CEL_ERR_BAD_MATRIX, "rows decrease after the sort (the reference's sort puts a "
                                            "last row of one entry among the row before it by column)");
        k_sub_int<<<grid_for(nnz), kThreads, 0, st>>>(nnz, co.p, cbase, co.p);
        CEL_HIP(hipGetLastError());
    }
    CEL_TRY(row_offsets_device(nrows, nnz, ro.p, rbase, 0, xadj.p));
    int rc = cel_gpu_mat_create_csr(A, nrows, ncols, nnz, xadj.p, co.p, vo.p, 0, 0, 0, CEL_MEM_DEVICE);
    if (rc < 0) return rc;
    (*A)->m.row_offset = row_offset;
    (*A)->m.col_offset = col_offset;
    CEL_HIP(hipStreamSynchronize(st));
    return CEL_OK;
}
