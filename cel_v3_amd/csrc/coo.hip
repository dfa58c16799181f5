// coo.hip -- COO ordering and COO -> CSR in HBM.
//
// cel_gpu_coo_sort replaces cel_sp_mat_format_sparse_31_sort
// (src/fspmt/cel_sp_mat_module.f90:518-585); cel_gpu_coo_row_offsets and
// cel_gpu_mat_create_coo_sorted replace cel_sp_mat_to_csr (:430-456) and the
// matrix build after it.
//
// The reference sorts on one CPU thread: a stable merge sort by row, then a
// loop that walks the row-sorted entries and merge-sorts by column each
// contiguous segment it closes.  Here:
//   stage 1  one stable LSD radix sort by row (rocPRIM through hipCUB), the
//            value and column travelling with the key (12 bytes), so no
//            entry is ever fetched from a random place;
//   groups   the contiguous groups stage 2 sorts: every row run
//            (CEL_SORT_ROW_COL), or the segments the reference's loop
//            closes, every entry after the last of them a group of its own
//            (CEL_SORT_CEL_COMPAT; transitions below), found with scans;
//   stage 2  a stable sort by column inside each group: one thread per
//            group of <= 32 entries (rank counting in LDS), one workgroup per
//            group of <= 4096 (bitonic sort of unique (column, position) keys
//            in LDS), and for anything longer one stable radix sort by
//            (group, column) over all entries.
// Every step is stable, and a stable sort's output is fixed by its keys and
// its input order, so the result is the reference's, entry for entry.
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
constexpr int kSmall = 32;     // groups up to this length: one thread each
constexpr int kMedium = 4096;  // up to this length: one workgroup each, in LDS

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

// What travels with the row key through the stage-1 radix sort: the value's
// bits and the local column (12 bytes).
struct Pay {
    uint32_t lo, hi, col;
};

__device__ inline double pay_val(const Pay& p)
{
    return __longlong_as_double((long long)(((unsigned long long)p.hi << 32) | p.lo));
}

#define CEL_GRID_STRIDE(i, n)                                                     \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
         i += (long long)gridDim.x * blockDim.x)

// Local 0-based row (the stage-1 key) and payload of every entry; range check.
__global__ void __launch_bounds__(kThreads)
k_pack(long long nnz, const int* __restrict__ rows, const int* __restrict__ cols,
       const double* __restrict__ vals, int rbase, int nrows, int cbase, int ncols,
       uint32_t* __restrict__ rk, Pay* __restrict__ pay, int* __restrict__ bad)
{
    int any = 0;
    CEL_GRID_STRIDE(i, nnz) {
        long long r = (long long)rows[i] - rbase;
        long long c = (long long)cols[i] - cbase;
        bool ok = r >= 0 && r < nrows && c >= 0 && c < ncols;
        any |= !ok;
        unsigned long long b = (unsigned long long)__double_as_longlong(vals[i]);
        rk[i] = ok ? (uint32_t)r : 0u;
        pay[i] = Pay{(uint32_t)b, (uint32_t)(b >> 32), ok ? (uint32_t)c : 0u};
    }
    if (any) atomicOr(bad, 1);
}

// Entries back to (rows, cols, vals) in the order they stand.
__global__ void __launch_bounds__(kThreads)
k_unpack(long long nnz, const uint32_t* __restrict__ rk, const Pay* __restrict__ pay, int rbase,
         int cbase, int* rows, int* cols, double* vals)
{
    CEL_GRID_STRIDE(p, nnz) {
        Pay q = pay[p];
        rows[p] = (int)rk[p] + rbase;
        cols[p] = (int)q.col + cbase;
        vals[p] = pay_val(q);
    }
}

// CEL_SORT_ROW_COL: every row run is one group; mark[q] = q closes a group.
__global__ void __launch_bounds__(kThreads)
k_row_ends(long long nnz, const uint32_t* __restrict__ r, int* __restrict__ mark)
{
    CEL_GRID_STRIDE(q, nnz) mark[q] = q == nnz - 1 || r[q + 1] != r[q];
}

// t[p]: is p a transition of the reference's column loop.
__global__ void __launch_bounds__(kThreads)
k_transitions(long long nnz, const uint32_t* __restrict__ r, long long row_offset,
              int* __restrict__ t)
{
    CEL_GRID_STRIDE(p, nnz) {
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
    CEL_GRID_STRIDE(p, nnz) if (t[p] && ord[p] == K) *pk = (int)p;
}

// CEL_SORT_CEL_COMPAT: mark[q] = 1 when q closes a segment the loop sorts,
// or lies after the last one (each such entry is a group of its own).
__global__ void __launch_bounds__(kThreads)
k_group_ends(long long nnz, const int* __restrict__ t, const int* __restrict__ ord, int K,
             const int* __restrict__ pk, int* __restrict__ mark)
{
    long long pK = *pk;
    long long endK = pK == nnz - 1 ? nnz - 1 : (pK > 1 ? pK - 1 : 0);
    CEL_GRID_STRIDE(q, nnz) {
        bool handled_next = q + 1 <= nnz - 2 && t[q + 1] && ord[q + 1] <= K;  // e(q+1) = q
        bool last = q == nnz - 1 && t[q] && ord[q] <= K;                     // e = nnz-1
        bool first = q == 0 && t[0] && ord[0] <= K;                          // e(0) = 0
        mark[q] = (handled_next || last || first || q > endK) ? 1 : 0;
    }
}

// head[q] = q opens a group.
__global__ void __launch_bounds__(kThreads)
k_heads(long long nnz, const int* __restrict__ mark, int* __restrict__ head)
{
    CEL_GRID_STRIDE(q, nnz) head[q] = q == 0 || mark[q - 1];
}

// starts[g] = first position of group g (ih = inclusive count of heads);
// starts[G] = nnz.
__global__ void __launch_bounds__(kThreads)
k_starts(long long nnz, const int* __restrict__ head, const int* __restrict__ ih,
         int* __restrict__ starts)
{
    CEL_GRID_STRIDE(q, nnz) {
        if (head[q]) starts[ih[q] - 1] = (int)q;
        if (q == nnz - 1) starts[ih[q]] = (int)nnz;
    }
}

// Groups of at most kSmall entries, one thread each: the rank of entry j is
// #{k : col_k < col_j} + #{k < j : col_k == col_j}, i.e. the stable order by
// column, and entry j goes to slot rank.  A workgroup takes 256 consecutive
// groups; when their entries fit (kTileCap) they are loaded into LDS and
// written back with coalesced accesses, otherwise each thread works on its
// group in global memory.  Longer groups are left in place here and listed
// for k_group_medium (order of the list is free: groups are independent).
constexpr int kTileCap = 2048;

__global__ void __launch_bounds__(kThreads)
k_group_small(int G, const int* __restrict__ starts, const uint32_t* __restrict__ rk,
              const Pay* __restrict__ pay, int rbase, int cbase, int* rows, int* cols,
              double* vals, int* __restrict__ longer, int* __restrict__ nlonger,
              int* __restrict__ maxlen)
{
    __shared__ int st[kThreads + 1];
    __shared__ uint32_t in_rk[kTileCap], out_rk[kTileCap];
    __shared__ Pay in_pay[kTileCap], out_pay[kTileCap];
    const int tid = threadIdx.x;
    for (long long g0 = (long long)blockIdx.x * kThreads; g0 < G; g0 += (long long)gridDim.x * kThreads) {
        const int ng = (int)min((long long)kThreads, G - g0);
        for (int i = tid; i <= ng; i += kThreads) st[i] = starts[g0 + i];
        __syncthreads();
        const int base = st[0];
        const int E = st[ng] - base;
        const int g = (int)g0 + tid;
        int s = 0, L = 0;
        if (tid < ng) {
            s = st[tid] - base;
            L = st[tid + 1] - st[tid];
            if (L > kSmall) {
                longer[atomicAdd(nlonger, 1)] = g;
                atomicMax(maxlen, L);
            }
        }
        if (E <= kTileCap) {
            const uint32_t* p32 = reinterpret_cast<const uint32_t*>(pay + base);
            uint32_t* l32 = reinterpret_cast<uint32_t*>(in_pay);
            for (int i = tid; i < 3 * E; i += kThreads) l32[i] = p32[i];
            for (int i = tid; i < E; i += kThreads) in_rk[i] = rk[base + i];
            __syncthreads();
            if (tid < ng) {
                if (L <= kSmall) {
                    for (int j = 0; j < L; ++j) {
                        const uint32_t cj = in_pay[s + j].col;
                        int r = 0;
                        for (int k = 0; k < L; ++k) {
                            const uint32_t ck = in_pay[s + k].col;
                            r += ck < cj || (ck == cj && k < j);
                        }
                        out_pay[s + r] = in_pay[s + j];
                        out_rk[s + r] = in_rk[s + j];
                    }
                } else {  // k_group_medium rewrites this group later
                    for (int j = 0; j < L; ++j) {
                        out_pay[s + j] = in_pay[s + j];
                        out_rk[s + j] = in_rk[s + j];
                    }
                }
            }
            __syncthreads();
            for (int i = tid; i < E; i += kThreads) {
                const Pay q = out_pay[i];
                rows[base + i] = (int)out_rk[i] + rbase;
                cols[base + i] = (int)q.col + cbase;
                vals[base + i] = pay_val(q);
            }
        } else if (tid < ng && L <= kSmall) {
            const int gs = base + s;
            for (int j = 0; j < L; ++j) {
                const uint32_t cj = pay[gs + j].col;
                int r = 0;
                for (int k = 0; k < L; ++k) {
                    const uint32_t ck = pay[gs + k].col;
                    r += ck < cj || (ck == cj && k < j);
                }
                const Pay q = pay[gs + j];
                rows[gs + r] = (int)rk[gs + j] + rbase;
                cols[gs + r] = (int)cj + cbase;
                vals[gs + r] = pay_val(q);
            }
        }
        __syncthreads();  // st and the tiles are reused by the next round
    }
}

// Groups of kSmall+1 .. kMedium entries, one workgroup each: bitonic sort
// in LDS of the unique keys (col << 32 | j), so equal columns keep their
// order; then entry j goes to its sorted slot.
__global__ void __launch_bounds__(kThreads)
k_group_medium(const int* __restrict__ list, const int* __restrict__ starts,
               const uint32_t* __restrict__ rk, const Pay* __restrict__ pay, int rbase, int cbase,
               int* rows, int* cols, double* vals)
{
    __shared__ unsigned long long key[kMedium];
    const int g = list[blockIdx.x];
    const int s = starts[g];
    const int L = starts[g + 1] - s;
    int P = 1;
    while (P < L) P <<= 1;
    for (int i = threadIdx.x; i < P; i += kThreads)
        key[i] = i < L ? ((unsigned long long)pay[s + i].col << 32) | (unsigned)i : ~0ULL;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += kThreads) {
                const int x = i ^ j;
                if (x > i) {
                    const unsigned long long a = key[i], b = key[x];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        key[i] = b;
                        key[x] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int m = threadIdx.x; m < L; m += kThreads) {
        const int j = (int)(key[m] & 0xffffffffULL);
        const Pay q = pay[s + j];
        rows[s + m] = (int)rk[s + j] + rbase;
        cols[s + m] = (int)q.col + cbase;
        vals[s + m] = pay_val(q);
    }
}

// Fallback for a group longer than kMedium: key = group << cb | col over all
// entries (group = ih - 1), sorted stably with positions as values.
__global__ void __launch_bounds__(kThreads)
k_group_keys(long long nnz, const int* __restrict__ ih, const Pay* __restrict__ pay, int cb,
             uint64_t* __restrict__ key, int* __restrict__ idx)
{
    CEL_GRID_STRIDE(p, nnz) {
        key[p] = ((uint64_t)(uint32_t)(ih[p] - 1) << cb) | pay[p].col;
        idx[p] = (int)p;
    }
}

__global__ void __launch_bounds__(kThreads)
k_gather_out(long long nnz, const int* __restrict__ perm, const uint32_t* __restrict__ rk,
             const Pay* __restrict__ pay, int rbase, int cbase, int* rows, int* cols, double* vals)
{
    CEL_GRID_STRIDE(p, nnz) {
        const int j = perm[p];
        const Pay q = pay[j];
        rows[p] = (int)rk[j] + rbase;
        cols[p] = (int)q.col + cbase;
        vals[p] = pay_val(q);
    }
}

// Rows of consecutive entries must not decrease for a CSR build.
__global__ void __launch_bounds__(kThreads)
k_rows_monotone(long long nnz, const int* __restrict__ rows, int* __restrict__ bad)
{
    int any = 0;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x + 1; p < nnz;
         p += (long long)gridDim.x * blockDim.x)
        any |= rows[p] < rows[p - 1];
    if (any) atomicOr(bad, 1);
}

// cel_sp_mat_to_csr: counts per row, then a scan.  cnt has nrows+1 slots;
// rows outside [0, nrows) are counted nowhere (flagged by k_rows_in_range).
__global__ void __launch_bounds__(kThreads)
k_row_count(long long nnz, const int* __restrict__ rows, int rbase, int nrows,
            int* __restrict__ cnt)
{
    CEL_GRID_STRIDE(p, nnz) {
        long long r = (long long)rows[p] - rbase;
        if (r >= 0 && r < nrows) atomicAdd(&cnt[r + 1], 1);
    }
}

__global__ void __launch_bounds__(kThreads)
k_rows_in_range(long long nnz, const int* __restrict__ rows, int rbase, int nrows,
                int* __restrict__ bad)
{
    int any = 0;
    CEL_GRID_STRIDE(p, nnz) {
        long long r = (long long)rows[p] - rbase;
        any |= r < 0 || r >= nrows;
    }
    if (any) atomicOr(bad, 1);
}

__global__ void __launch_bounds__(kThreads)
k_fill_int(long long n, int* __restrict__ a, int v)
{
    CEL_GRID_STRIDE(i, n) a[i] = v;
}

__global__ void __launch_bounds__(kThreads)
k_sub_int(long long n, const int* a, int v, int* out)  // may run in place
{
    CEL_GRID_STRIDE(i, n) out[i] = a[i] - v;
}

// The sort's device workspace, kept from call to call on the device that
// made it and grown when a larger sort needs more (a 117M-entry sort holds
// about 6 GB); allocating and freeing it per call cost about 5 ms of a
// 14 ms sort.  cel_gpu_finalize frees it.  Calls that overlap from several
// host threads share it and race.
struct SortBufs {
    DevBuf<uint32_t> rk0, rk1;
    DevBuf<Pay> pay0, pay1;
    DevBuf<int> a, b, c, starts, idx, perm, lists;
    DevBuf<uint64_t> k0, k1;
    DevBuf<unsigned char> temp;
    DevBuf<int> rows, cols, xadj, flag, cnt;  // staging and outputs of the entry points
    DevBuf<double> vals;
    int device = -1;
    void release()
    {
        rk0.release(); rk1.release(); pay0.release(); pay1.release();
        a.release(); b.release(); c.release(); starts.release(); idx.release(); perm.release();
        lists.release(); k0.release(); k1.release(); temp.release();
        rows.release(); cols.release(); xadj.release(); flag.release(); cnt.release();
        vals.release();
    }
};

// Never destroyed: a static destructor could run after the HIP runtime has
// gone at process exit.
SortBufs& sort_cache()
{
    static SortBufs* S = new SortBufs;
    if (S->device != rt().device) {
        S->release();
        S->device = rt().device;
    }
    return *S;
}

int temp_fit(SortBufs& S, size_t bytes)
{
    if (S.temp.n >= bytes && S.temp.p) return CEL_OK;
    CEL_HIP(hipStreamSynchronize(rt().stream));  // earlier passes may still read it
    return S.temp.alloc(bytes ? bytes : 1);
}

template <typename T>
int d2h_int(T* host, const T* dev)
{
    CEL_HIP(hipMemcpyAsync(host, dev, sizeof(T), hipMemcpyDeviceToHost, rt().stream));
    CEL_HIP(hipStreamSynchronize(rt().stream));
    return CEL_OK;
}

// Sorts (ri, ci, vi) into (ro, co, vo), device arrays; the outputs may be
// the inputs (the inputs are consumed by the first kernel).  Returns CEL_OK
// or CEL_ERR_REF_UNDEFINED (outputs then in the row-sorted order).
int sort_device(SortBufs& S, int nrows, int ncols, long long nnz, int rbase, int cbase,
                long long row_offset, int mode, const int* ri, const int* ci, const double* vi,
                int* ro, int* co, double* vo)
{
    hipStream_t st = rt().stream;
    const int g = grid_for(nnz);
    const int n = (int)nnz;
    const int rb = bits_for(nrows > 0 ? (unsigned long long)(nrows - 1) : 0ULL);
    CEL_TRY(S.rk0.reserve(nnz));
    CEL_TRY(S.rk1.reserve(nnz));
    CEL_TRY(S.pay0.reserve(nnz));
    CEL_TRY(S.pay1.reserve(nnz));
    CEL_TRY(S.lists.reserve(4));  // [0] bad, [1] K-th position, [2] longer count, [3] max length
    CEL_HIP(hipMemsetAsync(S.lists.p, 0, 4 * sizeof(int), st));
    k_pack<<<g, kThreads, 0, st>>>(nnz, ri, ci, vi, rbase, nrows, cbase, ncols, S.rk0.p, S.pay0.p,
                                    S.lists.p);
    CEL_HIP(hipGetLastError());
    int bad = 0;
    CEL_TRY(d2h_int(&bad, S.lists.p));
    if (bad) return fail(CEL_ERR_BAD_MATRIX, "row or column index out of range");

    // stage 1: stable by row, the payload carried along
    size_t bytes = 0;
    const int eb = rb > 0 ? rb : 1;
    CEL_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, S.rk0.p, S.rk1.p, S.pay0.p, S.pay1.p,
                                               n, 0, eb, st));
    CEL_TRY(temp_fit(S, bytes));
    CEL_HIP(hipcub::DeviceRadixSort::SortPairs(S.temp.p, bytes, S.rk0.p, S.rk1.p, S.pay0.p, S.pay1.p,
                                               n, 0, eb, st));
    auto unpack = [&]() -> int {
        k_unpack<<<g, kThreads, 0, st>>>(nnz, S.rk1.p, S.pay1.p, rbase, cbase, ro, co, vo);
        CEL_HIP(hipGetLastError());
        return CEL_OK;
    };

    // groups that are column-sorted
    CEL_TRY(S.a.reserve(nnz));
    CEL_TRY(S.b.reserve(nnz));
    CEL_TRY(S.c.reserve(nnz));
    int* mark = S.c.p;
    if (mode == CEL_SORT_ROW_COL) {
        k_row_ends<<<g, kThreads, 0, st>>>(nnz, S.rk1.p, mark);
        CEL_HIP(hipGetLastError());
    } else {
        const int K = nrows - 1;
        if (K <= 0) return unpack();  // the loop body never runs
        k_transitions<<<g, kThreads, 0, st>>>(nnz, S.rk1.p, row_offset, S.a.p);
        CEL_HIP(hipGetLastError());
        CEL_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, S.a.p, S.b.p, n, st));
        CEL_TRY(temp_fit(S, bytes));
        CEL_HIP(hipcub::DeviceScan::InclusiveSum(S.temp.p, bytes, S.a.p, S.b.p, n, st));
        int total = 0;
        CEL_TRY(d2h_int(&total, S.b.p + (nnz - 1)));
        if (total < K) {
            CEL_TRY(unpack());
            CEL_HIP(hipStreamSynchronize(st));
            return fail(CEL_ERR_REF_UNDEFINED,
                        "the reference's column loop needs nrows-1 row transitions and finds " +
                            std::to_string(total) + ": it reads past its arrays (empty rows)");
        }
        k_kth<<<g, kThreads, 0, st>>>(nnz, S.a.p, S.b.p, K, S.lists.p + 1);
        k_group_ends<<<g, kThreads, 0, st>>>(nnz, S.a.p, S.b.p, K, S.lists.p + 1, mark);
        CEL_HIP(hipGetLastError());
    }
    int* head = S.a.p;
    int* ih = S.b.p;
    k_heads<<<g, kThreads, 0, st>>>(nnz, mark, head);
    CEL_HIP(hipGetLastError());
    CEL_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, head, ih, n, st));
    CEL_TRY(temp_fit(S, bytes));
    CEL_HIP(hipcub::DeviceScan::InclusiveSum(S.temp.p, bytes, head, ih, n, st));
    CEL_TRY(S.starts.reserve(nnz + 1));
    k_starts<<<g, kThreads, 0, st>>>(nnz, head, ih, S.starts.p);
    CEL_HIP(hipGetLastError());
    int G = 0;
    CEL_TRY(d2h_int(&G, ih + (nnz - 1)));

    // stage 2: stable by column inside each group
    int* longer = mark;  // the group marks are spent; G <= nnz slots
    k_group_small<<<grid_for(G), kThreads, 0, st>>>(G, S.starts.p, S.rk1.p, S.pay1.p, rbase, cbase,
                                                     ro, co, vo, longer, S.lists.p + 2,
                                                     S.lists.p + 3);
    CEL_HIP(hipGetLastError());
    int nl[2] = {0, 0};
    CEL_HIP(hipMemcpyAsync(nl, S.lists.p + 2, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    CEL_HIP(hipStreamSynchronize(st));
    if (nl[0] == 0) return CEL_OK;
    if (nl[1] <= kMedium) {
        k_group_medium<<<nl[0], kThreads, 0, st>>>(longer, S.starts.p, S.rk1.p, S.pay1.p, rbase,
                                                    cbase, ro, co, vo);
        CEL_HIP(hipGetLastError());
        return CEL_OK;
    }
    // a group longer than the LDS sort holds: one stable radix sort by
    // (group, column) over everything
    const int cb = bits_for(ncols > 0 ? (unsigned long long)(ncols - 1) : 0ULL);
    const int sb = bits_for((unsigned long long)G);
    CEL_TRY(S.k0.reserve(nnz));
    CEL_TRY(S.k1.reserve(nnz));
    CEL_TRY(S.idx.reserve(nnz));
    CEL_TRY(S.perm.reserve(nnz));
    k_group_keys<<<g, kThreads, 0, st>>>(nnz, ih, S.pay1.p, cb, S.k0.p, S.idx.p);
    CEL_HIP(hipGetLastError());
    CEL_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, S.k0.p, S.k1.p, S.idx.p, S.perm.p, n,
                                               0, cb + sb, st));
    CEL_TRY(temp_fit(S, bytes));
    CEL_HIP(hipcub::DeviceRadixSort::SortPairs(S.temp.p, bytes, S.k0.p, S.k1.p, S.idx.p, S.perm.p,
                                               n, 0, cb + sb, st));
    k_gather_out<<<g, kThreads, 0, st>>>(nnz, S.perm.p, S.rk1.p, S.pay1.p, rbase, cbase, ro, co, vo);
    CEL_HIP(hipGetLastError());
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
    SortBufs& S = sort_cache();
    CEL_TRY(S.flag.reserve(1));
    CEL_TRY(S.cnt.reserve((size_t)nrows + 1));
    int* flag = S.flag.p;
    int* cnt = S.cnt.p;
    CEL_HIP(hipMemsetAsync(flag, 0, sizeof(int), st));
    k_fill_int<<<grid_for(nrows + 1LL), kThreads, 0, st>>>(nrows + 1LL, cnt, 0);
    k_fill_int<<<1, 1, 0, st>>>(1, cnt, out_base);  // xadj[0] = out_base carries through the scan
    if (nnz > 0) {
        k_rows_in_range<<<grid_for(nnz), kThreads, 0, st>>>(nnz, rows, rbase, nrows, flag);
        k_row_count<<<grid_for(nnz), kThreads, 0, st>>>(nnz, rows, rbase, nrows, cnt);
    }
    CEL_HIP(hipGetLastError());
    int bad = 0;
    CEL_HIP(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    size_t bytes = 0;
    CEL_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, cnt, xadj, nrows + 1, st));
    CEL_TRY(temp_fit(S, bytes));
    CEL_HIP(hipcub::DeviceScan::InclusiveSum(S.temp.p, bytes, cnt, xadj, nrows + 1, st));
    CEL_HIP(hipStreamSynchronize(st));
    if (bad) return fail(CEL_ERR_BAD_MATRIX, "row index out of range");
    return CEL_OK;
}

int launch_coo_to_csr(int nrows, int nnz, const int* rows0, int* xadj)
{
    return row_offsets_device(nrows, nnz, rows0, 0, 0, xadj);
}

void coo_release_workspace() { sort_cache().release(); }

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
    const int rbase = index_base + row_offset, cbase = index_base + col_offset;
    hipStream_t st = rt().stream;
    SortBufs& S = sort_cache();
    int rc;
    if (memspace == CEL_MEM_DEVICE) {
        rc = sort_device(S, nrows, ncols, nnz, rbase, cbase, row_offset, mode, rows, cols, vals, rows,
                         cols, vals);
    } else {
        CEL_TRY(S.rows.reserve(nnz));
        CEL_TRY(S.cols.reserve(nnz));
        CEL_TRY(S.vals.reserve(nnz));
        int* dr = S.rows.p;
        int* dc = S.cols.p;
        double* dv = S.vals.p;
        CEL_HIP(hipMemcpyAsync(dr, rows, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice, st));
        CEL_HIP(hipMemcpyAsync(dc, cols, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice, st));
        CEL_HIP(hipMemcpyAsync(dv, vals, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice, st));
        rc = sort_device(S, nrows, ncols, nnz, rbase, cbase, row_offset, mode, dr, dc, dv, dr, dc, dv);
        if (rc == CEL_OK || rc == CEL_ERR_REF_UNDEFINED) {
            CEL_HIP(hipMemcpyAsync(rows, dr, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToHost, st));
            CEL_HIP(hipMemcpyAsync(cols, dc, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToHost, st));
            CEL_HIP(hipMemcpyAsync(vals, dv, sizeof(double) * (size_t)nnz, hipMemcpyDeviceToHost, st));
        }
        CEL_HIP(hipStreamSynchronize(st));
    }
    CEL_HIP(hipStreamSynchronize(st));
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
    SortBufs& S = sort_cache();
    const int* r = rows;
    if (memspace == CEL_MEM_HOST && nnz > 0) {
        CEL_TRY(S.rows.reserve(nnz));
        CEL_HIP(hipMemcpyAsync(S.rows.p, rows, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice, st));
        r = S.rows.p;
    }
    int* x = xadj;
    if (memspace == CEL_MEM_HOST) {
        CEL_TRY(S.xadj.reserve((size_t)nrows + 1));
        x = S.xadj.p;
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
    SortBufs& S = sort_cache();
    DevBuf<int>& ro = S.rows;
    DevBuf<int>& co = S.cols;
    DevBuf<int>& xadj = S.xadj;
    DevBuf<int>& flag = S.flag;
    DevBuf<double>& vo = S.vals;
    CEL_TRY(ro.reserve(nnz > 0 ? nnz : 1));
    CEL_TRY(co.reserve(nnz > 0 ? nnz : 1));
    CEL_TRY(vo.reserve(nnz > 0 ? nnz : 1));
    CEL_TRY(xadj.reserve((size_t)nrows + 1));
    if (nnz > 0) {
        const int* ri = rows;
        const int* ci = cols;
        const double* vi = vals;
        if (memspace == CEL_MEM_HOST) {
            CEL_HIP(hipMemcpyAsync(ro.p, rows, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice, st));
            CEL_HIP(hipMemcpyAsync(co.p, cols, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice, st));
            CEL_HIP(hipMemcpyAsync(vo.p, vals, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice, st));
            ri = ro.p;
            ci = co.p;
            vi = vo.p;
        }
        CEL_TRY(sort_device(S, nrows, ncols, nnz, rbase, cbase, row_offset, mode, ri, ci, vi, ro.p,
                            co.p, vo.p));
        // a CSR build needs rows that never decrease; the reference's sort
        // can leave a last row of one entry inside the row before it
        CEL_TRY(flag.reserve(1));
        CEL_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), st));
        k_rows_monotone<<<grid_for(nnz), kThreads, 0, st>>>(nnz, ro.p, flag.p);
        CEL_HIP(hipGetLastError());
        int bad = 0;
        CEL_TRY(d2h_int(&bad, flag.p));
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
