// Block sparsity pattern built on the device.
//
// The reference (SystemAssembler.hh:251-465) sorts each element's block variables, appends rows
// bvars[0..i] to the bucket of column bvars[i], then deduplicates and sorts every bucket.  The same
// buckets are formed here: one thread per element counts its contributions per column (atomic
// integer adds), an exclusive scan gives the bucket offsets, a second pass writes the rows at
// atomically claimed positions inside each bucket, and one thread per column sorts and
// deduplicates its bucket.  The pattern is a set, so the arbitrary order of the atomic fills
// disappears in the sort: Ap/Ai equal the host construction's (and the reference's) exactly.
#include <hip/hip_runtime.h>

#include <climits>
#include <string>

#include "device.h"

namespace mfs {

#define PD_CHECK(x)                                                                          \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess)                                                                \
            throw Error(-4, std::string("HIP error: ") + hipGetErrorString(e_) + " at " +    \
                                __FILE__ + ":" + std::to_string(__LINE__));                 \
    } while (0)
#define PD_LAUNCH() PD_CHECK(hipGetLastError())

namespace {

constexpr int MAXNPE = 16;

// the element's nodes ascending in v[0..npe); entries npe.. hold INT_MAX (fixed-index network,
// so v stays in registers)
__device__ __forceinline__ bool load_sorted(const int32_t* __restrict__ en, int64_t e, int npe, int64_t nb,
                                            int32_t (&v)[MAXNPE]) {
    bool ok = true;
#pragma unroll
    for (int a = 0; a < MAXNPE; ++a) {
        v[a] = INT_MAX;
        if (a < npe) {
            v[a] = en[(int64_t)npe * e + a];
            ok = ok && v[a] >= 0 && v[a] < nb;
        }
    }
#pragma unroll
    for (int i = 1; i < MAXNPE; ++i)
#pragma unroll
        for (int j = MAXNPE - 1; j >= i; --j) {
            const int32_t lo = min(v[j - 1], v[j]), hi = max(v[j - 1], v[j]);
            v[j - 1] = lo;
            v[j] = hi;
        }
    return ok;
}

// bucket sizes: column v[i] receives i + 1 rows from the element
__global__ void k_pat_count(const int32_t* __restrict__ en, int64_t ne, int npe, int64_t nb,
                            unsigned long long* __restrict__ cnt, int* __restrict__ err) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    int32_t v[MAXNPE];
    if (!load_sorted(en, e, npe, nb, v)) {
        atomicOr(err, 1);
        return;
    }
#pragma unroll
    for (int i = 0; i < MAXNPE; ++i)
        if (i < npe) atomicAdd(&cnt[v[i]], (unsigned long long)(i + 1));
}

__global__ void k_pat_fill(const int32_t* __restrict__ en, int64_t ne, int npe, int64_t nb,
                           const int64_t* __restrict__ off, unsigned long long* __restrict__ cur,
                           int32_t* __restrict__ bucket) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    int32_t v[MAXNPE];
    if (!load_sorted(en, e, npe, nb, v)) return;
#pragma unroll
    for (int i = 0; i < MAXNPE; ++i)
        if (i < npe) {
            int32_t* dst = bucket + off[v[i]] + (int64_t)atomicAdd(&cur[v[i]], (unsigned long long)(i + 1));
#pragma unroll
            for (int k = 0; k <= i; ++k) dst[k] = v[k];
        }
}

// per column: sort the bucket (insertion sort, in place) and keep the unique rows in front
__global__ void k_pat_sort_unique(const int64_t* __restrict__ off, int64_t nb, int32_t* __restrict__ bucket,
                                  int64_t* __restrict__ ucnt) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    int32_t* b = bucket + off[j];
    const int64_t n = off[j + 1] - off[j];
    for (int64_t i = 1; i < n; ++i) {
        const int32_t x = b[i];
        int64_t k = i - 1;
        while (k >= 0 && b[k] > x) {
            b[k + 1] = b[k];
            --k;
        }
        b[k + 1] = x;
    }
    int64_t m = 0;
    for (int64_t i = 0; i < n; ++i)
        if (m == 0 || b[i] != b[m - 1]) b[m++] = b[i];
    ucnt[j] = m;
}

__global__ void k_pat_copy(const int64_t* __restrict__ off, const int64_t* __restrict__ Ap, int64_t nb,
                           const int32_t* __restrict__ bucket, int64_t* __restrict__ Ai) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    const int32_t* b = bucket + off[j];
    for (int64_t k = 0; k < Ap[j + 1] - Ap[j]; ++k) Ai[Ap[j] + k] = b[k];
}

// ---- exclusive scan of int64: out[0] = 0, out[i + 1] = in[0] + ... + in[i] ----------------
constexpr int SCAN_T = 256, SCAN_PER = 4, SCAN_B = SCAN_T * SCAN_PER;

// inclusive scan of each SCAN_B chunk into out[1..], chunk totals to sums
__global__ void k_scan_chunks(const int64_t* __restrict__ in, int64_t n, int64_t* __restrict__ out,
                              int64_t* __restrict__ sums) {
    __shared__ int64_t s[SCAN_T];
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * SCAN_B + (int64_t)t * SCAN_PER;
    int64_t v[SCAN_PER], acc = 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER; ++q) {
        v[q] = base + q < n ? in[base + q] : 0;
        acc += v[q];
        v[q] = acc;
    }
    s[t] = acc;
    __syncthreads();
    for (int d = 1; d < SCAN_T; d <<= 1) {
        const int64_t add = t >= d ? s[t - d] : 0;
        __syncthreads();
        s[t] += add;
        __syncthreads();
    }
    const int64_t before = t > 0 ? s[t - 1] : 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER; ++q)
        if (base + q < n) out[base + q + 1] = before + v[q];
    if (t == SCAN_T - 1) sums[blockIdx.x] = s[t];
    if (blockIdx.x == 0 && t == 0) out[0] = 0;
}

// add the exclusive prefix of the chunk totals to every chunk after the first
__global__ void k_scan_add(int64_t n, int64_t* __restrict__ out, const int64_t* __restrict__ chunk_prefix) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t c = i / SCAN_B;
    if (c > 0) out[i + 1] += chunk_prefix[c];
}

void scan_exclusive(const int64_t* in, int64_t n, int64_t* out, hipStream_t s) {
    if (n <= 0) {
        PD_CHECK(hipMemsetAsync(out, 0, sizeof(int64_t), s));
        return;
    }
    const int64_t nc = (n + SCAN_B - 1) / SCAN_B;
    DeviceArray sums, pre;
    sums.alloc(nc * sizeof(int64_t));
    k_scan_chunks<<<(unsigned)nc, SCAN_T, 0, s>>>(in, n, out, sums.as<int64_t>());
    PD_LAUNCH();
    if (nc > 1) {
        pre.alloc((nc + 1) * sizeof(int64_t));
        scan_exclusive(sums.as<int64_t>(), nc, pre.as<int64_t>(), s);
        k_scan_add<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, out, pre.as<int64_t>());
        PD_LAUNCH();
    }
    PD_CHECK(hipStreamSynchronize(s));   // the temporaries are freed on return
}

} // namespace

Pattern build_block_pattern_device(const int32_t* en_dev, int64_t ne, int npe, int64_t nb, void* stream,
                                   DeviceArray& Ap_d, DeviceArray& Ai_d) {
    if (npe <= 0 || npe > MAXNPE) throw Error(-1, "nodes_per_elem must be in 1..16");
    if (nb <= 0) throw Error(-1, "num_blocks must be positive");
    if (nb >= (int64_t(1) << 31)) throw Error(-6, "too many blocks for 32-bit rows");
    if (ne < 0 || (ne > 0 && !en_dev)) throw Error(-1, "element array");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    DeviceArray cnt, off, err, bucket;
    cnt.alloc(nb * sizeof(int64_t));
    off.alloc((nb + 1) * sizeof(int64_t));
    err.alloc(sizeof(int));
    PD_CHECK(hipMemsetAsync(cnt.p, 0, cnt.bytes, s));
    PD_CHECK(hipMemsetAsync(err.p, 0, sizeof(int), s));
    const unsigned ge = (unsigned)std::max<int64_t>((ne + 255) / 256, 1), gn = (unsigned)((nb + 255) / 256);
    if (ne > 0) {
        k_pat_count<<<ge, 256, 0, s>>>(en_dev, ne, npe, nb, cnt.as<unsigned long long>(), err.as<int>());
        PD_LAUNCH();
    }
    int herr = 0;
    PD_CHECK(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    PD_CHECK(hipStreamSynchronize(s));
    if (herr) throw Error(-1, "element node index out of range");
    scan_exclusive(cnt.as<int64_t>(), nb, off.as<int64_t>(), s);
    int64_t total = 0;
    PD_CHECK(hipMemcpy(&total, off.as<int64_t>() + nb, sizeof(int64_t), hipMemcpyDeviceToHost));
    bucket.alloc(std::max<int64_t>(total, 1) * sizeof(int32_t));
    PD_CHECK(hipMemsetAsync(cnt.p, 0, cnt.bytes, s));       // fill cursors
    if (ne > 0) {
        k_pat_fill<<<ge, 256, 0, s>>>(en_dev, ne, npe, nb, off.as<int64_t>(), cnt.as<unsigned long long>(),
                                      bucket.as<int32_t>());
        PD_LAUNCH();
    }
    k_pat_sort_unique<<<gn, 256, 0, s>>>(off.as<int64_t>(), nb, bucket.as<int32_t>(), cnt.as<int64_t>());
    PD_LAUNCH();
    Ap_d.alloc((nb + 1) * sizeof(int64_t));
    scan_exclusive(cnt.as<int64_t>(), nb, Ap_d.as<int64_t>(), s);
    int64_t nz = 0;
    PD_CHECK(hipMemcpy(&nz, Ap_d.as<int64_t>() + nb, sizeof(int64_t), hipMemcpyDeviceToHost));
    Ai_d.alloc(std::max<int64_t>(nz, 1) * sizeof(int64_t));
    k_pat_copy<<<gn, 256, 0, s>>>(off.as<int64_t>(), Ap_d.as<int64_t>(), nb, bucket.as<int32_t>(),
                                  Ai_d.as<int64_t>());
    PD_LAUNCH();
    Pattern P;
    P.n = nb;
    P.Ap.resize(nb + 1);
    P.Ai.resize(nz);
    PD_CHECK(hipMemcpyAsync(P.Ap.data(), Ap_d.p, (nb + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    if (nz) PD_CHECK(hipMemcpyAsync(P.Ai.data(), Ai_d.p, nz * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PD_CHECK(hipStreamSynchronize(s));
    return P;
}

} // namespace mfs
