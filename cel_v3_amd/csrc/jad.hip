// jad.hip -- CSR -> JAD conversion on the device, bit-exact with the
// reference converter (src/fspmtconverter/cel_sp_mat_jad_converter_module.f90:57-175).
//
// The reference sorts rows by length with dcsort, a stable descending
// counting sort: row i goes to position #{len > L_i} + #{j < i : len_j == L_i}.
// On the device:
//   1. row lengths, their min/max and a histogram        (one pass over xadj)
//   2. per 256-row block: count of each length, and each row's rank among
//      equal-length rows before it inside the block       (LDS, no atomics)
//   3. exclusive scan of the per-block counts over blocks, per length
//   4. perm[pos(i)] = i
//   5. diagonal scatter: entry d of permuted row k goes to begin[d] + k
// Everything that is O(n) or O(nnz) runs on the device; the host only turns
// the (mmax+1)-entry histogram into diagonal starts.
#include <algorithm>
#include <vector>

#include "cel_internal.hpp"
#include "device_common.hpp"

namespace cel {

constexpr int kJadBlock = 256;

__global__ void k_len_stats(int n, const int* __restrict__ xadj, int* __restrict__ mm /* min,max */)
{
    __shared__ int smin, smax;
    if (threadIdx.x == 0) {
        smin = 0x7fffffff;
        smax = 0;
    }
    __syncthreads();
    int lmin = 0x7fffffff, lmax = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int L = xadj[i + 1] - xadj[i];
        lmin = min(lmin, L);
        lmax = max(lmax, L);
    }
    atomicMin(&smin, lmin);
    atomicMax(&smax, lmax);
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicMin(&mm[0], smin);
        atomicMax(&mm[1], smax);
    }
}

// per-block counts cnt[bin * nblk + blk] and in-block stable rank of each row
__global__ void k_block_rank(int n, int ilo, int nbins, int nblk, const int* __restrict__ xadj,
                             int* __restrict__ cnt, int* __restrict__ rank)
{
    __shared__ int slen[kJadBlock];
    const int blk = blockIdx.x;
    const int i = blk * kJadBlock + threadIdx.x;
    const int L = i < n ? xadj[i + 1] - xadj[i] : -1;
    slen[threadIdx.x] = L;
    __syncthreads();
    if (i < n) {
        int r = 0;
        for (int t = 0; t < (int)threadIdx.x; ++t) r += (slen[t] == L);
        rank[i] = r;
        // the last row of each length in the block writes the block count
        bool last = true;
        for (int t = threadIdx.x + 1; t < kJadBlock; ++t)
            if (slen[t] == L) {
                last = false;
                break;
            }
        if (last) cnt[(size_t)(L - ilo) * nblk + blk] = r + 1;
    }
}

// exclusive scan of cnt over blocks for each bin; one block per bin.
// total[bin] receives the bin's row count.
__global__ void k_scan_bins(int nblk, int* __restrict__ cnt, int* __restrict__ total)
{
    __shared__ int sdata[kJadBlock];
    __shared__ int carry;
    int* c = cnt + (size_t)blockIdx.x * nblk;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nblk; base += kJadBlock) {
        const int j = base + threadIdx.x;
        const int v = j < nblk ? c[j] : 0;
        sdata[threadIdx.x] = v;
        __syncthreads();
        // inclusive Hillis-Steele scan in LDS
        for (int off = 1; off < kJadBlock; off <<= 1) {
            int add = threadIdx.x >= off ? sdata[threadIdx.x - off] : 0;
            __syncthreads();
            sdata[threadIdx.x] += add;
            __syncthreads();
        }
        if (j < nblk) c[j] = carry + sdata[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == kJadBlock - 1) carry += sdata[kJadBlock - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) total[blockIdx.x] = carry;
}

// perm[pos] = i, with pos = greater[bin] + blockoff + rank (0-based)
__global__ void k_place(int n, int ilo, int nblk, const int* __restrict__ xadj,
                        const int* __restrict__ cnt, const int* __restrict__ rank,
                        const int* __restrict__ greater, int* __restrict__ perm)
{
    const int i = blockIdx.x * kJadBlock + threadIdx.x;
    if (i >= n) return;
    const int bin = xadj[i + 1] - xadj[i] - ilo;
    const int pos = greater[bin] + cnt[(size_t)bin * nblk + blockIdx.x] + rank[i];
    perm[pos] = i;
}

// entry d of permuted row k -> begin[d] + k, for d < ndef and k < count[d]
template <typename VT>
__global__ void k_jad_scatter(int n, int ndef, const int* __restrict__ begin,
                              const int* __restrict__ perm, const int* __restrict__ xadj,
                              const int* __restrict__ cols, const VT* __restrict__ vals,
                              int* __restrict__ jcols, VT* __restrict__ jvals)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int row = perm[k];
    const int j0 = xadj[row];
    for (int d = 0; d < ndef; ++d) {
        const int b = begin[d];
        if (k >= begin[d + 1] - b) break;
        jcols[b + k] = cols[j0 + d];
        jvals[b + k] = vals[j0 + d];
    }
}

int jad_convert_device(Mat& A, int mode)
{
    CEL_TRY(ensure_init());
    if (mode != CEL_JAD_SPARSKIT && mode != CEL_JAD_CEL_COMPAT)
        return fail(CEL_ERR_ARG, "unknown JAD mode");
    hipStream_t s = rt().stream;
    const int n = A.n;
    A.jad_mode = -1;
    A.jad_pcols.release();
    if (n == 0) {
        A.jad_mmax = 0;
        A.jad_ilo = 0;
        A.jad_ndef = 0;
        A.jad_nnz = 0;
        A.h_jad_begin.assign(1, 0);
        CEL_TRY(A.jad_begin.alloc(1));
        CEL_HIP(hipMemcpy(A.jad_begin.p, A.h_jad_begin.data(), sizeof(int), hipMemcpyHostToDevice));
        A.jad_mode = mode;
        return CEL_OK;
    }
    // 1. min / max row length
    DevBuf<int> mm;
    CEL_TRY(mm.alloc(2));
    int init[2] = {0x7fffffff, 0};
    CEL_HIP(hipMemcpyAsync(mm.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    int g = std::min((n + kJadBlock - 1) / kJadBlock, rt().num_cus * 4);
    hipLaunchKernelGGL(k_len_stats, dim3(g), dim3(kJadBlock), 0, s, n, A.xadj.p, mm.p);
    CEL_HIP(hipGetLastError());
    int hmm[2];
    CEL_HIP(hipMemcpyAsync(hmm, mm.p, sizeof(hmm), hipMemcpyDeviceToHost, s));
    CEL_HIP(hipStreamSynchronize(s));
    // the reference starts ilo at nrow (:117), so ilo = min(nrow, min length)
    const int ilo = std::min(n, hmm[0]);
    const int ihi = hmm[1];
    const int nbins = ihi - ilo + 1;
    const int nblk = (n + kJadBlock - 1) / kJadBlock;
    if ((long long)nbins * nblk > (1LL << 28))
        return fail(CEL_ERR_ARG, "row-length range too wide for the device JAD converter");

    // 2-4. stable descending counting sort
    DevBuf<int> cnt, rank, total, greater;
    CEL_TRY(cnt.alloc((size_t)nbins * nblk));
    CEL_TRY(rank.alloc(n));
    CEL_TRY(total.alloc(nbins));
    CEL_TRY(greater.alloc(nbins));
    CEL_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * (size_t)nbins * nblk, s));
    hipLaunchKernelGGL(k_block_rank, dim3(nblk), dim3(kJadBlock), 0, s, n, ilo, nbins, nblk, A.xadj.p,
                       cnt.p, rank.p);
    hipLaunchKernelGGL(k_scan_bins, dim3(nbins), dim3(kJadBlock), 0, s, nblk, cnt.p, total.p);
    CEL_HIP(hipGetLastError());
    std::vector<int> htot(nbins), hgreater(nbins);
    CEL_HIP(hipMemcpyAsync(htot.data(), total.p, sizeof(int) * nbins, hipMemcpyDeviceToHost, s));
    CEL_HIP(hipStreamSynchronize(s));
    // greater[bin] = #{rows with length > ilo + bin}
    {
        int acc = 0;
        for (int b = nbins - 1; b >= 0; --b) {
            hgreater[b] = acc;
            acc += htot[b];
        }
    }
    CEL_HIP(hipMemcpyAsync(greater.p, hgreater.data(), sizeof(int) * nbins, hipMemcpyHostToDevice, s));
    CEL_TRY(A.jad_perm.alloc(n));
    hipLaunchKernelGGL(k_place, dim3(nblk), dim3(kJadBlock), 0, s, n, ilo, nblk, A.xadj.p, cnt.p,
                       rank.p, greater.p, A.jad_perm.p);
    CEL_HIP(hipGetLastError());

    // diagonal lengths (cel_sp_mat_jad_converter_module.f90:134-157)
    std::vector<long long> count(ihi + 1, 0);
    int ndef = ihi;
    if (mode == CEL_JAD_SPARSKIT || ilo == 0) {
        // #{len >= d}, d = 1..ihi
        for (int d = 1; d <= ihi; ++d) {
            long long c = 0;
            for (int b = std::max(d - ilo, 0); b < nbins; ++b) c += htot[b];
            count[d] = c;
        }
    } else {
        // dcsort leftovers: jad_begin(d) = #{len > ilo + d - 1} for d <= ihi-ilo+1,
        // uninitialised beyond
        ndef = std::min(nbins, ihi);
        for (int d = 1; d <= ndef; ++d) count[d] = hgreater[d - 1];
    }
    std::vector<int> hbegin(ihi + 1, -1);
    long long pos = 0;
    for (int d = 1; d <= ndef; ++d) {
        hbegin[d - 1] = (int)pos;
        pos += count[d];
    }
    // the start of diagonal ndef+1 is defined even when its count is not
    hbegin[ndef] = (int)pos;
    A.jad_mmax = ihi;
    A.jad_ilo = ilo;
    A.jad_ndef = ndef;
    A.jad_nnz = pos;
    A.h_jad_begin = hbegin;
    // device begin: undefined tail clamped to the written end so kernels stop
    std::vector<int> dbegin(hbegin);
    for (int d = ndef; d <= ihi; ++d) dbegin[d] = (int)pos;
    CEL_TRY(A.jad_begin.alloc(ihi + 1));
    CEL_HIP(hipMemcpyAsync(A.jad_begin.p, dbegin.data(), sizeof(int) * (ihi + 1), hipMemcpyHostToDevice, s));
    CEL_TRY(A.jad_cols.alloc(pos > 0 ? pos : 1));
    CEL_TRY(A.jad_vals.alloc(pos > 0 ? pos : 1));
    A.jad_vals32.release();
    // 5. scatter (begin is read by the kernel from device memory)
    const int gs = (n + kJadBlock - 1) / kJadBlock;
    hipLaunchKernelGGL(k_jad_scatter<double>, dim3(gs), dim3(kJadBlock), 0, s, n, ndef, A.jad_begin.p,
                       A.jad_perm.p, A.xadj.p, A.cols.p, A.vals.p, A.jad_cols.p, A.jad_vals.p);
    CEL_HIP(hipGetLastError());
    CEL_HIP(hipStreamSynchronize(s));
    A.jad_mode = mode;
    if (ndef != ihi)
        return fail(CEL_ERR_REF_UNDEFINED,
                    "CEL_COMPAT JAD: shortest row is non-empty, the reference's diagonal counts "
                    "beyond the defined ones are uninitialised");
    return CEL_OK;
}

// pinv[perm[k]] = k
__global__ void k_inv_perm(int n, const int* __restrict__ perm, int* __restrict__ pinv)
{
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x)
        pinv[perm[k]] = (int)k;
}

// out[e] = pinv[cols[e]]
__global__ void k_remap_cols(long long nnz, const int* __restrict__ cols,
                             const int* __restrict__ pinv, int* __restrict__ out)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nnz;
         e += (long long)gridDim.x * blockDim.x)
        out[e] = pinv[cols[e]];
}

int jad_prepare_permuted(Mat& A)
{
    if (A.jad_pcols.p || A.jad_nnz == 0) return CEL_OK;
    if (A.ncols != A.n) return fail(CEL_ERR_ARG, "the permuted JAD form needs a square block");
    hipStream_t s = rt().stream;
    DevBuf<int> pinv;
    CEL_TRY(pinv.alloc(A.n));
    CEL_TRY(A.jad_pcols.alloc(A.jad_nnz));
    const int g = std::max(1, std::min((A.n + kJadBlock - 1) / kJadBlock, rt().num_cus * 8));
    hipLaunchKernelGGL(k_inv_perm, dim3(g), dim3(kJadBlock), 0, s, A.n, A.jad_perm.p, pinv.p);
    const long long ge = std::max<long long>(
        1, std::min<long long>((A.jad_nnz + kJadBlock - 1) / kJadBlock, rt().num_cus * 8));
    hipLaunchKernelGGL(k_remap_cols, dim3((int)ge), dim3(kJadBlock), 0, s, A.jad_nnz, A.jad_cols.p,
                       pinv.p, A.jad_pcols.p);
    CEL_HIP(hipGetLastError());
    CEL_HIP(hipStreamSynchronize(s));  // pinv is released on return
    return CEL_OK;
}

}  // namespace cel
