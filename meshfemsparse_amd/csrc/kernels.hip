// gfx950 (MI355X) kernels of the per-Newton-step pipeline and the device-side driver objects.
//
//   assembly   : slot-map gather, fixed element order per 3x3 slot (HBM-bound)
//   injection  : plan gather of Ax blocks into supernode panels (+ sigma I / sigma B)
//   factor     : multifrontal supernodal LL^T over a static level schedule:
//                extend-add of children's update matrices, then per 64-column step
//                diagonal-tile POTRF (+ explicit tile inverse), TRSM as a GEMM with the inverse,
//                in-panel trailing update, and the Schur update U -= L21 L21^T;
//                all dense tile work runs through one batched FP64 MFMA (16x16x4) tile kernel
//   solves     : forward (children update vectors gathered in fixed order) and backward
//                (partial sums per row chunk, reduced in fixed order) supernodal sweeps
//   SpMV / IR  : block-CSR symmetric SpMV, deterministic norms, refinement driver
// Every reduction has a fixed order, so results are bitwise reproducible run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <mutex>

#include "../../include/meshfemsparse_amd.h"
#include "device.h"
#include "neohookean.h"

namespace mfs {

#define HIP_CHECK(x)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess)                                                                \
            throw Error(-4, std::string("HIP error: ") + hipGetErrorString(e_) + " at " +    \
                                __FILE__ + ":" + std::to_string(__LINE__));                 \
    } while (0)

static inline hipStream_t S_(void* s) { return reinterpret_cast<hipStream_t>(s); }
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
#define LAUNCH_CHECK() HIP_CHECK(hipGetLastError())

typedef double d4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------------------------------
// Device memory
// ------------------------------------------------------------------------------------------
DeviceArray::~DeviceArray() { release(); }
void DeviceArray::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}
void DeviceArray::alloc(size_t b) {
    if (p && bytes >= b) return;
    release();
    if (b == 0) b = 16;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
        p = nullptr;
        throw Error(-5, "hipMalloc of " + std::to_string(b) + " bytes failed: " + hipGetErrorString(e));
    }
    bytes = b;
}
void DeviceArray::upload_raw(const void* src, size_t b) {
    if (b) HIP_CHECK(hipMemcpy(p, src, b, hipMemcpyHostToDevice));
}

void KernelProfile::begin(int cls, void* stream, double work, int64_t items, double nbytes) {
    if (!on) return;
    work_.push_back({work, items});
    bytes_.push_back(nbytes);
    const size_t i = marks_.size() * 2;
    while (pool_.size() < i + 2) {
        hipEvent_t e;
        HIP_CHECK(hipEventCreate(&e));
        pool_.push_back(e);
    }
    HIP_CHECK(hipEventRecord((hipEvent_t)pool_[i], (hipStream_t)stream));
    marks_.push_back({cls, i});
}
void KernelProfile::end(void* stream) {
    if (!on || marks_.empty()) return;
    HIP_CHECK(hipEventRecord((hipEvent_t)pool_[marks_.back().second + 1], (hipStream_t)stream));
}
void KernelProfile::collect() {
    // MFS_PROFILE_DUMP=<path>: append one line per launch: class, tasks, work, milliseconds
    static const char* dump = getenv("MFS_PROFILE_DUMP");
    FILE* fp = dump ? fopen(dump, "a") : nullptr;
    for (size_t q = 0; q < marks_.size(); ++q) {
        auto [cls, i] = marks_[q];
        float m = 0;
        HIP_CHECK(hipEventSynchronize((hipEvent_t)pool_[i + 1]));
        HIP_CHECK(hipEventElapsedTime(&m, (hipEvent_t)pool_[i], (hipEvent_t)pool_[i + 1]));
        ms[cls] += m;
        count[cls] += 1;
        work[cls] += work_[q].first;
        bytes[cls] += bytes_[q];
        if (fp) fprintf(fp, "%d %lld %.6e %.6f\n", cls, (long long)work_[q].second, work_[q].first, m);
    }
    if (fp) fclose(fp);
    marks_.clear();
    work_.clear();
    bytes_.clear();
}
void KernelProfile::reset() {
    for (int c = 0; c < NCLASS; ++c) { ms[c] = 0; count[c] = 0; work[c] = 0; bytes[c] = 0; }
    marks_.clear();
    work_.clear();
    bytes_.clear();
}
KernelProfile::~KernelProfile() {
    for (void* e : pool_) (void)hipEventDestroy((hipEvent_t)e);
}

// ------------------------------------------------------------------------------------------
// Generic small kernels
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_set_int(int* p, int v) { *p = v; }
// Zero one tile per workgroup (thread: one row, 16 columns; coalesced column stores).
template <typename T>
__global__ __launch_bounds__(256) void k_fill_tiles(const FillTile* __restrict__ tiles, T* __restrict__ L) {
    const FillTile t = tiles[blockIdx.x];
    const int row = threadIdx.x & 63, cq = threadIdx.x >> 6;
    if (row >= t.M) return;
#pragma unroll
    for (int q = 0; q < 16; ++q)
        if (4 * q + cq < t.N) L[t.off + (int64_t)(4 * q + cq) * t.ld + row] = T(0);
}

template <typename T>
static void fill(T* p, int64_t n, T v, hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = (unsigned)std::min<int64_t>(nblk(n, 256), 8192);
    k_fill<T><<<g, 256, 0, s>>>(p, n, v);
    LAUNCH_CHECK();
}

// ------------------------------------------------------------------------------------------
// Assembly.  A wave is cut into 7 segments of 9 lanes; a segment owns one destination 3x3 slot
// and lane q of it the slot's entry q (column-major, row q % 3, column q / 3).  The segment walks
// the slot's contributions in ascending element order and every lane adds its entry of each
// contribution, so each entry receives the serial element-order sum of the reference, bit for
// bit, while the 9 lanes of a segment read one contribution's 9 values together (one 72-byte
// run of the element's local-upper Hessian) and write the slot's 9 values together.  A
// workgroup (4 waves, 28 consecutive slot positions) stages its contribution lists in LDS.
// Positions are grouped so that the slots of a workgroup have similar contribution counts
// (off-diagonal and diagonal slots apart), and consecutive workgroups run on one XCD, so the
// element Hessians the neighbouring slots share are fetched into one L2.
// ------------------------------------------------------------------------------------------
// Offset in He (in doubles) of entry (r, c) of contribution w: element w >> 5, local-upper block
// pair = (w >> 1) & 15 (6 off-diagonal blocks (a, b), a < b, then the 4 diagonal blocks),
// transposed if w & 1 (then entry (c, r) of the block is read).  `lo` holds the lane's in-block
// offsets for (r, c) and (c, r): off-diagonal blocks at lo.x / lo.y, diagonal blocks at lo.z / lo.w
// (UPPER78 only; r <= c there).  Selects only, so the loads of several contributions issue
// together.
struct AsmLane { int x, y, z, w; };

template <int LAYOUT>
__device__ __forceinline__ AsmLane asm_lane(int r, int c) {
    if constexpr (LAYOUT == 0) return AsmLane{12 * c + r, 12 * r + c, 0, 0};
    else return AsmLane{3 * c + r, 3 * r + c, (c * (c + 1)) / 2 + r, (r * (r + 1)) / 2 + c};
}

template <int LAYOUT>
__device__ __forceinline__ uint32_t he_local(uint32_t w, const AsmLane& lo) {
    const uint32_t pair = (w >> 1) & 15u;
    const bool tr = (w & 1u) != 0;
    if constexpr (LAYOUT == 0) {
        // 36 b + 3 a of block (a, b), 8 bits per pair: (0,1) (0,2) (0,3) (1,2) (1,3) (2,3) (0,0)
        // (1,1) | (2,2) (3,3)
        constexpr uint64_t K0 = 36ull | 72ull << 8 | 108ull << 16 | 75ull << 24 | 111ull << 32 | 114ull << 40 |
                                0ull << 48 | 39ull << 56;
        constexpr uint64_t K1 = 78ull | 117ull << 8;
        const uint32_t ab = (uint32_t)((pair < 8 ? K0 >> (8 * pair) : K1 >> (8 * (pair - 8))) & 0xffu);
        return ab + (uint32_t)(tr ? lo.y : lo.x);
    } else {
        return pair < 6 ? 9 * pair + (uint32_t)(tr ? lo.y : lo.x) : 18 + 6 * pair + (uint32_t)(tr ? lo.w : lo.z);
    }
}

// He element stride in doubles
template <int LAYOUT> constexpr uint32_t he_stride() { return LAYOUT == 0 ? 144u : 78u; }

constexpr int kAsmSeg = 9;                     // lanes per slot
constexpr int kAsmSlotsPerWave = 7;            // 63 of 64 lanes
constexpr int kAsmSlotsPerWG = 4 * kAsmSlotsPerWave;
constexpr int kAsmLds = 2048;                  // staged contribution entries per workgroup

// workgroup b of a launch of n workgroups runs on XCD b % 8 (round-robin dispatch); it takes
// chunk (b % 8) * ceil(n / 8) + b / 8, so every XCD works through one contiguous range
__device__ __forceinline__ int64_t xcd_chunk(int64_t b, int64_t n) {
    const int64_t per = (n + 7) / 8;
    return (b % 8) * per + b / 8;
}

// WIDE: He spans 4 GiB or more (64-bit element offsets); otherwise 32-bit byte offsets from He.
template <int LAYOUT, bool WIDE>
__global__ __launch_bounds__(256) void k_assemble(const int32_t* __restrict__ seg,
                                                  const uint32_t* __restrict__ contrib,
                                                  const int32_t* __restrict__ slot_of_pos,
                                                  const uint8_t* __restrict__ is_diag,
                                                  const double* __restrict__ He,
                                                  const int64_t* __restrict__ ncbase,
                                                  const int32_t* __restrict__ ncstride,
                                                  double* __restrict__ Ax, int64_t npos) {
    __shared__ uint32_t lc[kAsmLds];
    const int64_t chunk = xcd_chunk(blockIdx.x, gridDim.x);
    const int64_t p0 = chunk * kAsmSlotsPerWG;
    if (p0 >= npos) return;                    // the whole workgroup leaves together
    const int64_t pend = min(p0 + (int64_t)kAsmSlotsPerWG, npos);
    const int32_t cb = seg[p0], ce = seg[pend];
    const bool staged = ce - cb <= kAsmLds;
    if (staged)
        for (int32_t k = threadIdx.x; k < ce - cb; k += 256) lc[k] = contrib[cb + k];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int sg = lane / kAsmSeg, q = lane % kAsmSeg;
    const int r = q % 3, c = q / 3;
    const int64_t p = p0 + (threadIdx.x >> 6) * kAsmSlotsPerWave + sg;
    if (sg >= kAsmSlotsPerWave || p >= pend) return;
    const bool diag = is_diag[p] != 0;
    if (diag && r > c) return;                 // strict lower part of a diagonal block: untouched
    const int64_t s = slot_of_pos[p];
    // contiguous blocks: 9 values per slot; non-contiguous (scalar upper CSC): column c of the
    // slot's block starts c * stride + c (c + 1) / 2 after the slot's base
    const int64_t dst = ncbase ? ncbase[s] + (int64_t)c * ncstride[s] + (c * (c + 1)) / 2 + r : 9 * s + q;
    double v = Ax[dst];
    const int32_t b = seg[p], e_ = seg[p + 1];
    const AsmLane lo = asm_lane<LAYOUT>(r, c);
    auto load = [&](uint32_t w) -> double {
        if constexpr (WIDE)
            return He[(int64_t)(w >> 5) * he_stride<LAYOUT>() + he_local<LAYOUT>(w, lo)];
        else
            return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(He) +
                                                    ((w >> 5) * he_stride<LAYOUT>() + he_local<LAYOUT>(w, lo)) * 8u);
    };
    // batches of 8 contributions: their loads issue together (past the end a batch re-reads the
    // last contribution and skips it), the values are added in ascending element order
    constexpr int B = 8;                       // 4: the same time; 16: 15 % slower
    if (staged) {
        for (int32_t k = b; k < e_; k += B) {
            uint32_t w[B];
#pragma unroll
            for (int i = 0; i < B; ++i) w[i] = lc[min(k + i, e_ - 1) - cb];
            double x[B];
#pragma unroll
            for (int i = 0; i < B; ++i) x[i] = load(w[i]);
#pragma unroll
            for (int i = 0; i < B; ++i)
                if (k + i < e_) v += x[i];
        }
    } else {
        for (int32_t k = b; k < e_; k += B) {
            uint32_t w[B];
#pragma unroll
            for (int i = 0; i < B; ++i) w[i] = contrib[min(k + i, e_ - 1)];
            double x[B];
#pragma unroll
            for (int i = 0; i < B; ++i) x[i] = load(w[i]);
#pragma unroll
            for (int i = 0; i < B; ++i)
                if (k + i < e_) v += x[i];
        }
    }
    Ax[dst] = v;
}

// ------------------------------------------------------------------------------------------
// Element kernel: neo-Hookean Hessian (12x12 or its local-upper 78 values) and gradient of every
// linear tet, one thread per element (the element math is shared with the host generator).
// Gradient assembly: one thread per node sums its elements' contributions in ascending element
// order (the order of a serial element loop), accumulating into g.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_neohookean(const double* __restrict__ Xr, const double* __restrict__ X,
                                                    const int32_t* __restrict__ tets, int64_t ne,
                                                    const double* __restrict__ Ee, const double* __restrict__ nue,
                                                    double E, double nu, double* __restrict__ He, int layout,
                                                    double* __restrict__ g12) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= ne) return;
    double xr[4][3], x[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int64_t v = tets[4 * e + a];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            xr[a][r] = Xr[3 * v + r];
            x[a][r] = X[3 * v + r];
        }
    }
    NHState st;
    nh_state(xr, x, Ee ? Ee[e] : E, nue ? nue[e] : nu, st);
    if (g12) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int i = 0; i < 3; ++i) g12[12 * e + 3 * a + i] = nh_grad(st, a, i);
    }
    if (!He) return;
    if (layout == 0) {
        double* h = He + 144 * e;
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int i = 0; i < 3; ++i) h[(3 * b + j) * 12 + 3 * a + i] = nh_hess(st, a, i, b, j);
    } else {
        // local-upper packing: off-diagonal (a, b), a < b, as column-major 3x3, then the upper
        // triangles of the diagonal blocks
        const int A[6] = {0, 0, 0, 1, 1, 2}, B[6] = {1, 2, 3, 2, 3, 3};
        double* h = He + 78 * e;
#pragma unroll
        for (int q = 0; q < 6; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int r = 0; r < 3; ++r) h[9 * q + 3 * c + r] = nh_hess(st, A[q], r, B[q], c);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int r = 0; r <= c; ++r) h[54 + 6 * a + (c * (c + 1)) / 2 + r] = nh_hess(st, a, r, a, c);
    }
}

__global__ __launch_bounds__(256) void k_grad_gather(const int64_t* __restrict__ nptr, const int64_t* __restrict__ nent,
                                                     const double* __restrict__ g12, double* __restrict__ g,
                                                     int64_t nnodes) {
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= nnodes) return;
    double s0 = g[3 * v], s1 = g[3 * v + 1], s2 = g[3 * v + 2];
    for (int64_t q = nptr[v]; q < nptr[v + 1]; ++q) {
        const int64_t w = nent[q];                   // (element << 2) | local node
        const double* src = g12 + 12 * (w >> 2) + 3 * (w & 3);
        s0 += src[0];
        s1 += src[1];
        s2 += src[2];
    }
    g[3 * v] = s0;
    g[3 * v + 1] = s1;
    g[3 * v + 2] = s2;
}

AssemblerDev::AssemblerDev(const Pattern& P, const int32_t* en, int64_t ne, int npe) {
    SlotMap M = build_slot_map(P, en, ne, npe);
    ap_ = P.Ap;
    diag_last_ = true;
    for (int64_t j = 0; j < P.n; ++j)
        diag_last_ = diag_last_ && P.Ap[j + 1] > P.Ap[j] && P.Ai[P.Ap[j + 1] - 1] == j;
    nslots_ = M.nslots;
    ne_ = ne;
    npe_ = npe;
    nnodes_ = P.n;
    seg_.upload(M.seg);
    contrib_.upload(M.contrib);
    diag_.upload(M.is_diag);
    pslot_.upload(M.slot_of_pos);
    // connectivity and node -> (element, local node) lists in ascending element order
    std::vector<int32_t> conn(en, en + ne * npe);
    tets_.upload(conn);
    std::vector<int64_t> ptr(P.n + 1, 0);
    for (int64_t q = 0; q < ne * npe; ++q) ptr[en[q] + 1]++;
    for (int64_t v = 0; v < P.n; ++v) ptr[v + 1] += ptr[v];
    std::vector<int64_t> ent(ptr[P.n]);
    std::vector<int64_t> bk(ptr.begin(), ptr.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < npe; ++a) ent[bk[en[e * npe + a]]++] = (e << 2) | a;
    nptr_.upload(ptr);
    nent_.upload(ent);
}

void AssemblerDev::neohookean(const double* Xr, const double* X, const double* Ee, const double* nue, double E,
                              double nu, double* He, int layout, double* g12, void* stream) {
    if (npe_ != 4) throw Error(-6, "the element kernel handles linear tetrahedra (4 nodes)");
    if (layout != 0 && layout != 1) throw Error(-1, "unknown element Hessian layout");
    if (ne_ == 0) return;
    k_neohookean<<<nblk(ne_, 256), 256, 0, S_(stream)>>>(Xr, X, tets_.as<int32_t>(), ne_, Ee, nue, E, nu, He, layout,
                                                         g12);
    LAUNCH_CHECK();
}

void AssemblerDev::assemble_gradient(const double* g12, double* g, void* stream) {
    if (npe_ != 4) throw Error(-6, "gradient assembly handles 4-node elements");
    if (nnodes_ == 0) return;
    k_grad_gather<<<nblk(nnodes_, 256), 256, 0, S_(stream)>>>(nptr_.as<int64_t>(), nent_.as<int64_t>(), g12, g,
                                                              nnodes_);
    LAUNCH_CHECK();
}

// Non-contiguous layout (scalar upper CSC): block column j starts at 9 Ap[j] - 3 j; its scalar
// column c holds 3 (ne - 1) + c + 1 entries (3 rows per off-diagonal block, then the diagonal
// block's rows 0..c), so block idx of the column starts at base = start + 3 idx in column 0 and
// column c of it c * 3 (ne - 1) + c (c + 1) / 2 further on.
void AssemblerDev::build_noncontiguous_map() {
    if (!diag_last_) throw Error(-1, "the non-contiguous layout needs every diagonal block in the pattern");
    const int64_t n = (int64_t)ap_.size() - 1;
    std::vector<int64_t> base(nslots_);
    std::vector<int32_t> stride(nslots_);
    for (int64_t j = 0; j < n; ++j) {
        const int64_t ncol = ap_[j + 1] - ap_[j], s0 = 9 * ap_[j] - 3 * j;
        if (3 * (ncol - 1) > INT32_MAX) throw Error(-1, "column too long for the non-contiguous map");
        for (int64_t k = ap_[j]; k < ap_[j + 1]; ++k) {
            base[k] = s0 + 3 * (k - ap_[j]);
            stride[k] = (int32_t)(3 * (ncol - 1));
        }
    }
    ncbase_.upload(base);
    ncstride_.upload(stride);
}

void AssemblerDev::assemble(const double* He, int layout, double* Ax, void* stream, bool contiguous) {
    if (layout != 0 && layout != 1) throw Error(-1, "unknown element Hessian layout");
    if (!contiguous && ncbase_.bytes == 0 && nslots_ > 0) build_noncontiguous_map();
    if (nslots_ == 0) return;
    const int64_t* nb = contiguous ? nullptr : ncbase_.as<int64_t>();
    const int32_t* ns = contiguous ? nullptr : ncstride_.as<int32_t>();
    const int64_t nwg = (nslots_ + kAsmSlotsPerWG - 1) / kAsmSlotsPerWG;
    const unsigned grid = (unsigned)(8 * ((nwg + 7) / 8));       // whole rounds of the 8 XCDs
    const bool wide = (uint64_t)ne_ * (layout == 0 ? 144u : 78u) * 8u >= (1ull << 32);
    auto go = [&](auto kern) {
        kern<<<grid, 256, 0, S_(stream)>>>(seg_.as<int32_t>(), contrib_.as<uint32_t>(), pslot_.as<int32_t>(),
                                           diag_.as<uint8_t>(), He, nb, ns, Ax, nslots_);
    };
    if (layout == 0) wide ? go(k_assemble<0, true>) : go(k_assemble<0, false>);
    else wide ? go(k_assemble<1, true>) : go(k_assemble<1, false>);
    LAUNCH_CHECK();
}

void AssemblerDev::zero(double* Ax, void* stream) { fill(Ax, 9 * nslots_, 0.0, S_(stream)); }

// ------------------------------------------------------------------------------------------
// Value injection: L[dst] = block (as-is / transposed / lower of diagonal) + shift, optionally
// scaled symmetrically, s_i s_j (A + shift)_ij, and converted to the factor type.
// ------------------------------------------------------------------------------------------
struct PlanEnt {
    int64_t dstmode;
    int32_t src;
    int32_t ld;
    int32_t rb, cb;     // factor-order block row / column of the lower-triangle block
};

// cmask (NULL: none) flags masked components per factor-order block (bit c): pinned scalars of a
// partially pinned block and the padding of a scalar matrix; their rows and columns become the
// identity, so the free scalars factor as the reduced matrix.
// The values of one injection entry: put(r, c, v) for every stored entry (r, c) of the 3x3 block
// (a diagonal block's lower part), v = (A + sigma B or sigma I), symmetrically scaled, identity on
// masked components, rounded to T once.
template <typename T, typename Put>
__device__ __forceinline__ void inject_block(const PlanEnt& pe, const double* __restrict__ Ax,
                                             const double* __restrict__ Bx, double sigma,
                                             const double* __restrict__ scale,
                                             const uint8_t* __restrict__ cmask, Put put) {
    const int mode = (int)(pe.dstmode & 3);
    const double* a = Ax + 9 * (int64_t)pe.src;
    const double* bb = Bx ? Bx + 9 * (int64_t)pe.src : nullptr;
    const int mr = cmask ? cmask[pe.rb] : 0, mc = cmask ? cmask[pe.cb] : 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            if (mode == 2 && r < c) continue;
            const int srcidx = (mode == 0) ? 3 * c + r : 3 * r + c;
            double v = a[srcidx];
            if (bb) v += sigma * bb[srcidx];
            else if (mode == 2 && r == c) v += sigma;
            if (scale) v *= scale[3 * (int64_t)pe.rb + r] * scale[3 * (int64_t)pe.cb + c];
            if (((mr >> r) | (mc >> c)) & 1) v = (mode == 2 && r == c) ? 1.0 : 0.0;
            put(r, c, (T)v);
        }
}

template <typename T>
__global__ __launch_bounds__(256) void k_inject(const PlanEnt* __restrict__ plan, int64_t n,
                                                const double* __restrict__ Ax,
                                                const double* __restrict__ Bx, double sigma,
                                                const double* __restrict__ scale,
                                                const uint8_t* __restrict__ cmask,
                                                T* __restrict__ L) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const PlanEnt pe = plan[q];
    T* d = L + (pe.dstmode >> 2);
    inject_block<T>(pe, Ax, Bx, sigma, scale, cmask, [&](int r, int c, T v) { d[c * (int64_t)pe.ld + r] = v; });
}

// s_k = 1 / sqrt((A + shift)_kk) in factor order; a non-positive diagonal is a pivot failure
// (masked components: s = 1)
__global__ void k_diag_scale(const int32_t* __restrict__ diag_src, int64_t nblocks,
                             const double* __restrict__ Ax, const double* __restrict__ Bx,
                             double sigma, const uint8_t* __restrict__ cmask, double* __restrict__ scale,
                             int* __restrict__ status) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= nblocks) return;
    const int64_t src = 9 * (int64_t)diag_src[k];
    const int m = cmask ? cmask[k] : 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double d = 0.0;
        if (src >= 0) d = Ax[src + 4 * c] + (Bx ? sigma * Bx[src + 4 * c] : sigma);
        else if (!Bx) d = sigma;
        if ((m >> c) & 1) d = 1.0;
        if (!(d > 0.0)) {
            atomicMin(status, (int)(3 * k + c));
            d = 1.0;
        }
        scale[3 * k + c] = 1.0 / sqrt(d);
    }
}

// ------------------------------------------------------------------------------------------
// Extend-add: child's U (lower 3x3 blocks of column block kj) into the parent front.
// ------------------------------------------------------------------------------------------
// Four tasks per workgroup, one per wave: a CU holds a bounded number of workgroups, so
// one-wave workgroups would leave most of its wave slots empty when tasks are short.
template <typename T>
__global__ __launch_bounds__(256) void k_extend(const ExtendTask* __restrict__ tasks, int64_t ntasks,
                                                const int32_t* __restrict__ rel,
                                                T* __restrict__ U, T* __restrict__ L) {
    // lane = one scalar row of the column block, so every load and store of a wave covers
    // consecutive addresses (child rows are contiguous; parent rows follow the increasing
    // relative indices); each parent entry still receives exactly one add per launch
    const int64_t ti = 4 * (int64_t)blockIdx.x + (threadIdx.x >> 6);
    if (ti >= ntasks) return;                  // the whole wave leaves together
    const int lane = threadIdx.x & 63;
    const ExtendTask t = tasks[ti];
    const int32_t* relc = rel + t.rel_off;
    const int kj = t.col_block;
    const int64_t Dc = 3 * (int64_t)t.dc;
    const int tj = relc[kj];
    const T* src = U + t.u_child + (3 * (int64_t)kj) * Dc;
    T* dst;
    int64_t ldd, shift;
    if (tj < t.wp) {
        dst = L + t.p_panel + (int64_t)(3 * tj) * t.ldp;
        ldd = t.ldp;
        shift = 0;
    } else {
        const int64_t Dp = 3 * (int64_t)t.dp;
        dst = U + t.u_parent + (int64_t)(3 * (tj - t.wp)) * Dp;
        ldd = Dp;
        shift = 3 * (int64_t)t.wp;
    }
    // UN row chunks per pass: every index and value load of the pass is issued before its
    // stores (the child and parent matrices share one arena, so the compiler may not move loads
    // of the next rows above these stores by itself)
    constexpr int UN = 4;
    const int64_t row1 = min(Dc, (int64_t)t.row0 + kExtendRows);
    for (int64_t i0 = t.row0 + (int64_t)lane; i0 < row1; i0 += 64 * UN) {
        int64_t d[UN];
        int r[UN];
        bool ok[UN], diag[UN];
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            const int64_t i = i0 + 64 * u;
            ok[u] = i < row1;
            const int ki = ok[u] ? (int)(i / 3) : kj;
            r[u] = (int)(i - 3 * (int64_t)ki);
            diag[u] = ki == kj;     // diagonal block: its upper part (r < c) is not stored
            d[u] = 3 * (int64_t)relc[ki] + r[u] - shift;
        }
        T sv[UN][3], dv[UN][3];
#pragma unroll
        for (int u = 0; u < UN; ++u)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const bool on = ok[u] && (!diag[u] || r[u] >= c);
                sv[u][c] = on ? src[c * Dc + i0 + 64 * u] : T(0);
                dv[u][c] = on ? dst[c * ldd + d[u]] : T(0);
            }
#pragma unroll
        for (int u = 0; u < UN; ++u)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                if (ok[u] && (!diag[u] || r[u] >= c)) dst[c * ldd + d[u]] = dv[u][c] + sv[u][c];
    }
}

// ------------------------------------------------------------------------------------------
// Diagonal tile POTRF (kb <= 64) + explicit inverse X = L^{-1} of the tile's Cholesky factor,
// kept for the TRSM of this step and for both solve sweeps.  The tile sits in LDS (a[c][r]) and
// is factored in 8 steps of 8 columns; x[c][r] holds X(r, c) once final and, before that, the
// partial sum S(r, c) = sum_k L(r, k) X(k, c) of the block forward substitution
// X_bj = -X_bb sum_{k=j}^{b-1} L_bk X_kj.  Step b (columns b0 .. b0+7):
//   phase 1  wave 0: every lane factors the 8x8 diagonal block redundantly in registers
//            (left-looking Cholesky and X_bb, no cross-lane traffic on the critical path), then
//            lane i solves panel row b0+8+i, L(r, b) = A(r, b) X_bb^T;
//            waves 1-3: the rank-8 update from block b-1 of block columns b+1 .. 7 (deferred from
//            step b-1) and S += L(:, b-1) X(b-1, :) for the block rows b .. 7;
//   phase 2  wave 0: lane c < b0 turns S(b0.., c) into X(b0.., c) = -X_bb S; lane 0 stores
//            L_bb and X_bb; waves 1-3: the rank-8 update from block b of block column b+1.
// A barrier ends each phase; the updates of waves 1-3 run as 16x16 FP64 MFMA tiles (K = 8) masked
// to their regions, round robin over the three waves.  A
// non-positive or NaN pivot records the first failing column and continues with 1 so no NaN
// spreads; the host turns it into a pivot failure.
// ------------------------------------------------------------------------------------------
constexpr int TS = 64;   // diagonal tile size (scalars)
constexpr int PB = 8;    // column step of the tile factorization

// One 16x16 tile of a rank-8 update inside the diagonal tile, on FP64 MFMA (16x16x4, two steps):
//   INV = false: A(r, c) -= sum_k L(r, k0 + k) L(c, k0 + k)   (C = a, B(k, c) = a[k0 + k][c])
//   INV = true:  S(r, c) += sum_k L(r, k0 + k) X(k0 + k, c)   (C = x, B(k, c) = x[c][k0 + k])
// for the rows R0 .. R0+15 and columns C0 .. C0+15 with r >= rlo, clo <= c < chi and, if `lower`,
// r >= c; other entries of the tile are neither read nor written.  Operand layouts of
// v_mfma_f64_16x16x4f64: A(i, k) in lane i + 16 k, B(k, j) in lane j + 16 k, D(i, j) in lane
// j + 16 (i % 4), register i / 4.
template <bool INV>
__device__ __forceinline__ void tile_rank8(double (*__restrict__ a)[TS + 1], double (*__restrict__ x)[TS + 1],
                                           int k0, int R0, int C0, int rlo, int clo, int chi, bool lower) {
    const int lane = threadIdx.x & 63, i = lane & 15, kq = lane >> 4;
    double (*cm)[TS + 1] = INV ? x : a;
    const int c = C0 + i;
    d4 acc;
    bool ok[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = R0 + 4 * q + kq;
        ok[q] = r >= rlo && c >= clo && c < chi && (!lower || r >= c);
        acc[q] = ok[q] ? cm[c][r] : 0.0;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int k = k0 + 4 * h + kq;
        const double av = a[k][R0 + i];
        const double bv = INV ? x[C0 + i][k] : a[k][C0 + i];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(INV ? av : -av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (ok[q]) cm[c][R0 + 4 * q + kq] = acc[q];
}

// One tile (256 threads, a and x in LDS).  On return the factor is stored to the panel, the
// inverse to inv[t.slot], and x still holds X in LDS (all threads past the final barrier).
// The tile is factored and inverted in double whatever the storage type T: in the FP32 forms the
// diagonal-block pivots, the tile's own updates and its inverse are FP64 computations on the FP32
// input, rounded to FP32 only when stored (a and x are double arrays for every T).
// Shared tiles: with `counter`, `count` workgroups factor the same tile; each counts itself in
// after its tile load has completed, and only the last to arrive stores the factor, the inverse
// and a pivot failure (so no workgroup can see a stored factor in place of its input), then
// resets the counter for the next factorization.  No workgroup waits for another.
template <typename T>
__device__ __forceinline__ void potrf_tile(const PotrfTask& t, T* __restrict__ L, T* __restrict__ inv,
                                           int* __restrict__ status, double (*__restrict__ a)[TS + 1],
                                           double (*__restrict__ x)[TS + 1], int* __restrict__ counter = nullptr,
                                           int count = 1, int* __restrict__ s_flag = nullptr) {
    const int tid = threadIdx.x, kb = t.kb, wave = tid >> 6;
    const int64_t ld = t.ld;
    T* D = L + t.diag_off;
    {
        // every load of a thread is issued before the first use (one memory round trip); entries
        // outside the kb x kb tile read a clamped in-tile address and are replaced by the identity
        constexpr int PT = TS * TS / 256;
        T v[PT];
#pragma unroll
        for (int q = 0; q < PT; ++q) {
            const int idx = tid + 256 * q, r = idx & 63, c = idx >> 6;
            v[q] = D[(int64_t)min(c, kb - 1) * ld + min(r, kb - 1)];
        }
#pragma unroll
        for (int q = 0; q < PT; ++q) {
            const int idx = tid + 256 * q, r = idx & 63, c = idx >> 6;
            a[c][r] = (r < kb && c < kb) ? ((r >= c) ? (double)v[q] : 0.0) : ((r == c) ? 1.0 : 0.0);
            x[c][r] = 0.0;
        }
    }
    __syncthreads();    // every load of the tile has returned
    // count in now; the returned value is only needed at the end, so the atomic's round trip
    // overlaps the factorization
    int arrived = 0;
    if (counter && tid == 0) arrived = atomicAdd(counter, 1);
    int fail = INT_MAX;    // first failing column (wave 0, lane 0)
    constexpr int NBK = TS / PB;
#pragma unroll 1
    for (int bs = 0; bs < NBK; ++bs) {
        const int b0 = PB * bs;
        double lb[PB][PB], xb[PB][PB];  // L_bb, X_bb (lower triangles), wave 0
        if (wave == 0) {
#pragma unroll
            for (int r = 0; r < PB; ++r)
#pragma unroll
                for (int c = 0; c <= r; ++c) lb[r][c] = a[b0 + c][b0 + r];
            double rd[PB];               // 1 / L(j, j)
#pragma unroll
            for (int j = 0; j < PB; ++j) {
                double d = lb[j][j];
#pragma unroll
                for (int k = 0; k < j; ++k) d -= lb[j][k] * lb[j][k];
                if (!(d > 0.0)) {
                    fail = min(fail, t.col0 + b0 + j);
                    d = 1.0;
                }
                const double sd = sqrt(d);
                lb[j][j] = sd;
                rd[j] = 1.0 / sd;
#pragma unroll
                for (int i = j + 1; i < PB; ++i) {
                    double v = lb[i][j];
#pragma unroll
                    for (int k = 0; k < j; ++k) v -= lb[i][k] * lb[j][k];
                    lb[i][j] = v * rd[j];
                }
            }
#pragma unroll
            for (int c = 0; c < PB; ++c) {
                xb[c][c] = rd[c];
#pragma unroll
                for (int i = c + 1; i < PB; ++i) {
                    double v = 0.0;
#pragma unroll
                    for (int k = c; k < i; ++k) v += lb[i][k] * xb[k][c];
                    xb[i][c] = -v * rd[i];
                }
            }
            const int r = b0 + PB + tid;
            if (r < TS) {
                double ar[PB];
#pragma unroll
                for (int k = 0; k < PB; ++k) ar[k] = a[b0 + k][r];
#pragma unroll
                for (int c = 0; c < PB; ++c) {
                    double v = 0.0;
#pragma unroll
                    for (int k = 0; k <= c; ++k) v += ar[k] * xb[c][k];
                    a[b0 + c][r] = v;
                }
            }
        } else if (bs > 0) {
            // deferred from step bs-1 (panel k0): the trailing update of columns >= b0 + 8 and
            // S += L(:, k0..) X(k0.., :) for rows >= b0, columns < b0; 16x16 MFMA tiles (waves
            // 1-3 round robin), anchored at multiples of 16 and masked to those regions
            const int k0 = b0 - PB;
            const int t0 = (b0 + PB) / 16;                   // first tile row/column of the trailing part
            const int m = 4 - t0, na = m * (m + 1) / 2;
            const int s0 = b0 / 16, ns = (4 - s0) * ((b0 + 15) / 16);
            for (int q = wave - 1; q < na + ns; q += 3) {
                if (q < na) {
                    int tj = t0, e = q;
                    while (e >= 4 - tj) { e -= 4 - tj; ++tj; }
                    tile_rank8<false>(a, x, k0, 16 * (tj + e), 16 * tj, b0 + PB, b0 + PB, TS, true);
                } else {
                    const int e = q - na, nsc = (b0 + 15) / 16;
                    tile_rank8<true>(a, x, k0, 16 * (s0 + e / nsc), 16 * (e % nsc), b0, 0, b0, false);
                }
            }
        }
        __syncthreads();
        if (wave == 0) {
            if (tid < b0) {
                const int c = tid;
                double sv[PB];
#pragma unroll
                for (int i = 0; i < PB; ++i) sv[i] = x[c][b0 + i];
#pragma unroll
                for (int i = 0; i < PB; ++i) {
                    double v = 0.0;
#pragma unroll
                    for (int k = 0; k <= i; ++k) v += xb[i][k] * sv[k];
                    x[c][b0 + i] = -v;
                }
            }
            if (tid == 0) {
#pragma unroll
                for (int r = 0; r < PB; ++r)
#pragma unroll
                    for (int c = 0; c <= r; ++c) {
                        a[b0 + c][b0 + r] = lb[r][c];
                        x[b0 + c][b0 + r] = xb[r][c];
                    }
            }
        } else if (bs + 1 < NBK) {
            // lookahead: block column bs+1 (columns b0 + 8 .. b0 + 15) from this step's panel
            const int tc = (b0 + PB) / 16;
            for (int tr = tc + wave - 1; tr < 4; tr += 3)
                tile_rank8<false>(a, x, b0, 16 * tr, 16 * tc, b0 + PB, b0 + PB, b0 + 2 * PB, true);
        }
        __syncthreads();
    }
    if (counter) {
        if (tid == 0) {
            const bool last = arrived == count - 1;
            if (last) atomicExch(counter, 0);
            *s_flag = last;
        }
        __syncthreads();
        if (*s_flag == 0) return;
    }
    if (tid == 0 && fail != INT_MAX) atomicMin(status, fail);
    T* out = inv + (int64_t)t.slot * TS * TS;
#pragma unroll
    for (int q = 0; q < TS * TS / 256; ++q) {
        const int idx = tid + 256 * q, r = idx & 63, c = idx >> 6;
        const double lv = a[c][r], xv = x[c][r];
        if (r < kb && c < kb && r >= c) D[(int64_t)c * ld + r] = (T)lv;
        out[idx] = (r < kb && c < kb) ? (T)xv : T(0);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_potrf(const PotrfTask* __restrict__ tasks,
                                               T* __restrict__ L, T* __restrict__ inv,
                                               int* __restrict__ status) {
    __shared__ double a[TS][TS + 1];    // a[c][r] = tile(r, c) -> L(r, c)
    __shared__ double x[TS][TS + 1];    // x[c][r] = S(r, c) -> X(r, c)
    potrf_tile<T>(tasks[blockIdx.x], L, inv, status, a, x);
}

// ------------------------------------------------------------------------------------------
// Batched tile GEMM, C = alpha A B^T + beta C, TM x TN output tile per 256-thread workgroup
// (2x2 waves), MFMA 16x16x4 (FP64 or FP32), K staged through double-buffered LDS in chunks of 16.
// Operand layouts on gfx950 (measured): A[i][k] in lane i + 16k, B[k][j] in lane j + 16k for both;
// D[i][j] in lane j + 16 (i % 4), register i / 4 for f64_16x16x4 (rows interleaved across the four
// 16-lane groups), and in lane j + 16 (i / 4), register i % 4 for f32_16x16x4.
// ------------------------------------------------------------------------------------------
typedef float f4 __attribute__((ext_vector_type(4)));

// Workgroups are dispatched to the 8 XCDs round-robin by workgroup id.  Mode 1 gives XCD x the
// x-th contiguous run of tasks; mode 2 deals runs of 8 consecutive tasks round-robin (tasks
// beyond the last full round of 64 keep their position), so each run of 8 tiles shares one L2
// while the XCDs still draw evenly from the whole launch; mode 3 deals runs of 64 (an 8 x 8
// block of tiles in the task order of the schedule, whose A and B panels then share one L2;
// tasks beyond the last full round of 512 keep their position).
__device__ __forceinline__ int xcd_task(int b, int n, int mode) {
    if (mode == 1) {
        const int x = b & 7, j = b >> 3, q = n >> 3, r = n & 7;
        return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + j;
    }
    if (mode == 2 && b < (n & ~63)) {
        const int x = b & 7, j = b >> 3;
        return (((j >> 3) << 3) + x) * 8 + (j & 7);
    }
    if (mode == 3 && b < (n & ~511)) {
        const int x = b & 7, j = b >> 3;
        return (((j >> 6) << 3) + x) * 64 + (j & 63);
    }
    return b;
}

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef d4 acc_t;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <> struct Mfma<float> {
    typedef f4 acc_t;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// C (+)= alpha * acc for one wave's FM x FN fragments: per fragment row, all loads of C are issued
// before any store (one memory round trip per row instead of one per element).
template <typename T, int FM, int FN>
__device__ __forceinline__ void gemm_store_acc(const typename Mfma<T>::acc_t (&acc)[FM][FN], T* __restrict__ C,
                                               int64_t ldc, int M, int N, int wm, int wn, int lane,
                                               T alpha, bool beta0) {
    const int col_l = lane & 15;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        T cv[FN][4];
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = wn + 16 * j + col_l, row = wm + 16 * i + Mfma<T>::row(lane, r);
                cv[j][r] = (!beta0 && row < M && col < N) ? C[col * ldc + row] : T(0);
            }
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = wn + 16 * j + col_l, row = wm + 16 * i + Mfma<T>::row(lane, r);
                if (row < M && col < N) C[col * ldc + row] = cv[j][r] + alpha * acc[i][j][r];
            }
    }
}

// C (+)= alpha * acc for one wave's WM x WN quadrant through LDS: the quadrant goes to the wave's
// own LDS region in NP parts of WN / NP columns, each written in the MFMA accumulator layout and
// read back column by column, so the global loads and stores of C are whole column segments of WM
// rows (the accumulator layout itself touches 16 columns per instruction), and every load of a
// part is issued before its first store.  Same arithmetic as gemm_store_acc (cv + alpha * acc).
// `region` holds WM * WN / NP elements; rows are XOR-swizzled by the column so that both the
// layout writes and the column reads spread over the LDS banks.  Called by all waves together,
// after the K loop's last barrier (the operand buffers that hold the regions are no longer read).
// Raw buffer loads (resource word 3 for gfx9: 32-bit data format, no swizzle, stride 0).
constexpr int kBufWord3 = 0x00020000;
template <typename T> __device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <> __device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <> __device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

//
// Schur-update tiles (beta = 0, C in the update arena) may also carry their children's update-matrix
// blocks that land in them (`npull` PullEnt entries, child order).  The wave's values become
// v = 0 + alpha * acc in registers, and each child's entries are added in child order: the same
// additions, in the same order, as the extend-add passes into U after the Schur update.  Per entry
// the wave's rows and columns are mapped to the child's scalar rows and columns (-1: none) through
// LDS (`map`, WM + WN ints of the wave's epilogue region), then every fragment element with a
// child entry in the child's lower triangle loads it (up to four fragment rows of one fragment
// column, 16 loads, in flight).
template <typename T, int FM, int FN>
__device__ __forceinline__ void apply_pulls(typename Mfma<T>::acc_t (&acc)[FM][FN], T alpha,
                                            int* __restrict__ map, const PullEnt* __restrict__ pe, int npull,
                                            const int32_t* __restrict__ rel, const T* __restrict__ U,
                                            int wm, int wn, int lane) {
    constexpr int WM = 16 * FM, WN = 16 * FN;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] = T(0) + alpha * acc[i][j][r];
    int* rmap = map;
    int* cmap = map + WM;
    for (int e = 0; e < npull; ++e) {
        const PullEnt pu = pe[e];
        const int32_t* rl = rel + pu.rel_off;
        // wave-relative scalar row / column of each block's first scalar (lane b: blocks b, b + 64)
        int rb[2], cb[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int b = lane + 64 * h;
            rb[h] = b < pu.nib ? 3 * rl[pu.ib0 + b] - pu.base_r - wm : -(1 << 20);
            cb[h] = b < pu.njb ? 3 * rl[pu.jb0 + b] - pu.base_c - wn : -(1 << 20);
        }
        __syncthreads();     // the previous entry's map reads are done (npull is uniform per tile)
        for (int q = lane; q < WM + WN; q += 64) map[q] = -1;
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int ib = pu.ib0 + lane + 64 * h, jb = pu.jb0 + lane + 64 * h;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                if (rb[h] + r >= 0 && rb[h] + r < WM) rmap[rb[h] + r] = 3 * ib + r;
                if (cb[h] + r >= 0 && cb[h] + r < WN) cmap[cb[h] + r] = 3 * jb + r;
            }
        }
        __syncthreads();
        // buffer loads based at the entry's first child column (32-bit offsets within its 3 njb
        // columns), and an element outside the child's lower triangle reads the out-of-range
        // offset `nrec`, which returns +0 (v holds 0 + alpha * acc, never -0, so adding +0 leaves
        // it unchanged bit for bit)
        const int c0s = 3 * pu.jb0;
        const int nrec = 3 * pu.njb * pu.dc3 * (int)sizeof(T);
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(U + pu.u_child + (int64_t)c0s * pu.dc3), (short)0, nrec, kBufWord3);
        constexpr int FB = FM < 4 ? FM : 4;   // fragment rows of loads in flight (a whole column)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int cj = cmap[16 * j + (lane & 15)];
            const int cbase = (cj - c0s) * pu.dc3;
#pragma unroll
            for (int i0 = 0; i0 < FM; i0 += FB) {
                T v[FB][4];
#pragma unroll
                for (int i = 0; i < FB; ++i)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int ri = rmap[16 * (i0 + i) + Mfma<T>::row(lane, r)];
                        // child lower triangle: scalar row >= scalar column (cj >= 0 then implies ri >= 0)
                        const int off = (cj >= 0 && ri >= cj) ? (cbase + ri) * (int)sizeof(T) : nrec;
                        v[i][r] = buf_load<T>(rs, off, 0);
                    }
#pragma unroll
                for (int i = 0; i < FB; ++i)
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc[i0 + i][j][r] += v[i][r];
            }
        }
    }
    __syncthreads();         // map reads done before the epilogue reuses the region
}

// The children's update-matrix blocks that land in a 64 x 64 tile of a parent's panel, added to
// the tile in child order: the additions, in the order, of the extend-add passes into the panel.
// The tile's starting values are rebuilt in LDS from its injection entries (zero elsewhere, as the
// clear and the injection make the tiles without lists), so L is only written.  Thread (row = tid % 64,
// columns 4 q + tid / 64) then holds 16 entries of the tile; per child the block rows and columns
// are mapped to the child's scalar rows and columns through LDS, and every entry with a child
// entry in the child's lower triangle loads it and adds it.
template <typename T>
__global__ __launch_bounds__(256) void k_panel_pull(const PanelPullTask* __restrict__ tasks,
                                                    const PullEnt* __restrict__ pulls,
                                                    const int32_t* __restrict__ rel,
                                                    const TileInj* __restrict__ tinj,
                                                    const PlanEnt* __restrict__ plan,
                                                    const double* __restrict__ Ax,
                                                    const double* __restrict__ Bx, double sigma,
                                                    const double* __restrict__ scale,
                                                    const uint8_t* __restrict__ cmask,
                                                    T* __restrict__ L, const T* __restrict__ U) {
    __shared__ T tile[64 * 64];
    __shared__ int map[128];   // rows, then columns: child scalar index or -1
    const PanelPullTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x, row = tid & 63, cq = tid >> 6;
    T* C = L + t.c_off;
#pragma unroll
    for (int q = 0; q < 16; ++q) tile[(4 * q + cq) * 64 + row] = T(0);
    __syncthreads();
    for (int i = tid; i < t.ninj; i += 256) {     // distinct blocks: no two threads share an entry
        const TileInj ti = tinj[t.inj_begin + i];
        const PlanEnt pe = plan[ti.plan];
        inject_block<T>(pe, Ax, Bx, sigma, scale, cmask, [&](int r, int c, T v) {
            const int lr = ti.lr + r, lc = ti.lc + c;
            if (lr >= 0 && lr < 64 && lc >= 0 && lc < 64) tile[lc * 64 + lr] = v;
        });
    }
    __syncthreads();
    T val[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) val[q] = tile[(4 * q + cq) * 64 + row];
    for (int e = 0; e < t.npull; ++e) {
        const PullEnt pu = pulls[t.pull_begin + e];
        const int32_t* rl = rel + pu.rel_off;
        __syncthreads();                       // the previous entry's map reads are done
        if (tid < 128) map[tid] = -1;
        __syncthreads();
        if (tid < pu.nib) {
            const int ib = pu.ib0 + tid, rb = 3 * rl[ib] - pu.base_r;
#pragma unroll
            for (int r = 0; r < 3; ++r)
                if (rb + r >= 0 && rb + r < 64) map[rb + r] = 3 * ib + r;
        } else if (tid >= 64 && tid - 64 < pu.njb) {
            const int jb = pu.jb0 + tid - 64, cb = 3 * rl[jb] - pu.base_c;
#pragma unroll
            for (int r = 0; r < 3; ++r)
                if (cb + r >= 0 && cb + r < 64) map[64 + cb + r] = 3 * jb + r;
        }
        __syncthreads();
        const int ri = map[row];
        // buffer loads based at the entry's first child column (32-bit offsets within its 3 njb
        // columns; an element without a child entry reads the out-of-range offset nrec)
        const int c0s = 3 * pu.jb0;
        const int nrec = 3 * pu.njb * pu.dc3 * (int)sizeof(T);
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(U + pu.u_child + (int64_t)c0s * pu.dc3), (short)0, nrec, kBufWord3);
        int cj[16];
        T v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            cj[q] = map[64 + 4 * q + cq];
            // child lower triangle: scalar row >= scalar column (cj >= 0 then implies ri >= 0)
            const int off = (cj[q] >= 0 && ri >= cj[q]) ? ((cj[q] - c0s) * pu.dc3 + ri) * (int)sizeof(T) : nrec;
            v[q] = buf_load<T>(rs, off, 0);
        }
        // only entries with a child entry are added (an injected -0 stays -0, as in the passes)
#pragma unroll
        for (int q = 0; q < 16; ++q)
            if (cj[q] >= 0 && ri >= cj[q]) val[q] += v[q];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int col = 4 * q + cq;
        if (row < t.M && col < t.N) C[(int64_t)col * t.ldc + row] = val[q];
    }
}

// (`final_vals`: acc already holds the values to store, as after apply_pulls)
template <typename T, int FM, int FN, int NP>
__device__ __forceinline__ void gemm_store_lds(const typename Mfma<T>::acc_t (&acc)[FM][FN], T* __restrict__ C,
                                               int64_t ldc, int M, int N, int wm, int wn, int lane,
                                               T alpha, bool beta0, T* __restrict__ region,
                                               bool final_vals = false) {
    constexpr int FP = FN / NP, WM = 16 * FM, WC = 16 * FP, PER = WM * WC / 64;   // per part
    static_assert(FN % NP == 0 && WM % 32 == 0 && PER * 64 == WM * WC, "epilogue shape");
#pragma unroll
    for (int h = 0; h < NP; ++h) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int jj = 0; jj < FP; ++jj)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * i + Mfma<T>::row(lane, r), col = 16 * jj + (lane & 15);
                    region[col * WM + (row ^ (col & 31))] = acc[i][h * FP + jj][r];
                }
        __syncthreads();
        T cv[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = lane + 64 * q, row = e % WM, col = e / WM;
            const int grow = wm + row, gcol = wn + h * WC + col;
            cv[q] = (!beta0 && grow < M && gcol < N) ? C[gcol * ldc + grow] : T(0);
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = lane + 64 * q, row = e % WM, col = e / WM;
            const int grow = wm + row, gcol = wn + h * WC + col;
            const T rv = region[col * WM + (row ^ (col & 31))];
            if (grow < M && gcol < N) C[gcol * ldc + grow] = final_vals ? rv : cv[q] + alpha * rv;
        }
        __syncthreads();
    }
}


// One output tile of a task (256 threads; As/Bs: double-buffered operand chunks in LDS, rows padded
// by PAD elements).  The operand loads are raw buffer loads: per K chunk one buffer resource based at
// column k0 of the operand (uniform, built by scalar instructions), a per-thread 32-bit row offset
// and a uniform per-load column offset, so the loop carries no per-load vector address arithmetic.
// Rows beyond M (N) read row 0 and only feed output rows (columns) that are never stored, so full
// chunks are staged unmasked; the one partial chunk clamps k to K - 1 and zeroes k >= K at the
// LDS store.  Rows are padded (gemm_pad) so the 4 k rows one MFMA step reads fall in different
// banks.  The loads of the next chunk stay in flight under the current chunk's MFMAs.
// LDS row padding of the operand chunks (elements): FP32 rows by 16 words and the FP64 128-row
// tile by 16 doubles, so the four k rows one MFMA step reads (16 lanes each) fall in different
// bank halves: conflict-free fragment reads (FP64: 0.2785 -> 0.2764 s per 64^3 step; the FP64
// 64-row tile is unchanged with or without it and keeps its smaller LDS footprint)
template <typename T, int TM> constexpr int gemm_pad() { return sizeof(T) == 4 ? 16 : (TM == 128 ? 16 : 0); }
template <typename T, int TM, int TN, int KC, int PAD = gemm_pad<T, TM>()>
__device__ __forceinline__ void gemm_tile(const GemmTask& t, T* __restrict__ L, T* __restrict__ U,
                                            const T* __restrict__ inv, T (*__restrict__ As)[KC][TM + PAD],
                                            T (*__restrict__ Bs)[KC][TN + PAD],
                                            const PullEnt* __restrict__ pulls = nullptr,
                                            const int32_t* __restrict__ rel = nullptr) {
    typedef typename Mfma<T>::acc_t acc_t;
    constexpr int WM = TM / 2, WN = TN / 2, FM = WM / 16, FN = WN / 16;
    constexpr int AST = 256 / TM, BST = 256 / TN;
    constexpr int AL = KC / AST, BL = KC / BST;
    const T* A = L + t.a_off;
    const T* B = (t.flags & GT_B_IN_INV) ? inv + t.b_off : L + t.b_off;
    T* C = (t.flags & GT_C_IN_U) ? U + t.c_off : L + t.c_off;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * WM, wn = (wave >> 1) * WN;
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
    const int ar = tid % TM, ak = tid / TM, br = tid % TN, bk = tid / TN;
    const int lda = t.lda, ldb = t.ldb, K = t.K;
    const int arow = ar < t.M ? ar : 0, brow = br < t.N ? br : 0;
    const int va = (arow + ak * lda) * (int)sizeof(T), vb = (brow + bk * ldb) * (int)sizeof(T);
    const int ea = KC * lda * (int)sizeof(T), eb = KC * ldb * (int)sizeof(T);   // chunk extents
    T ra[AL], rb[BL];
    auto gload = [&](int k0) {
        const __amdgpu_buffer_rsrc_t rA =
            __builtin_amdgcn_make_buffer_rsrc((void*)(A + (int64_t)k0 * lda), (short)0, ea, kBufWord3);
        const __amdgpu_buffer_rsrc_t rB =
            __builtin_amdgcn_make_buffer_rsrc((void*)(B + (int64_t)k0 * ldb), (short)0, eb, kBufWord3);
        if (k0 + KC <= K) {
#pragma unroll
            for (int q = 0; q < AL; ++q) ra[q] = buf_load<T>(rA, va, AST * q * lda * (int)sizeof(T));
#pragma unroll
            for (int q = 0; q < BL; ++q) rb[q] = buf_load<T>(rB, vb, BST * q * ldb * (int)sizeof(T));
        } else {
            const int kl = K - 1 - k0;
#pragma unroll
            for (int q = 0; q < AL; ++q)
                ra[q] = buf_load<T>(rA, (arow + min(ak + AST * q, kl) * lda) * (int)sizeof(T), 0);
#pragma unroll
            for (int q = 0; q < BL; ++q)
                rb[q] = buf_load<T>(rB, (brow + min(bk + BST * q, kl) * ldb) * (int)sizeof(T), 0);
        }
    };
    auto sstore = [&](int buf, int k0) {
        if (k0 + KC <= K) {
#pragma unroll
            for (int q = 0; q < AL; ++q) As[buf][ak + AST * q][ar] = ra[q];
#pragma unroll
            for (int q = 0; q < BL; ++q) Bs[buf][bk + BST * q][br] = rb[q];
        } else {
#pragma unroll
            for (int q = 0; q < AL; ++q) As[buf][ak + AST * q][ar] = (k0 + ak + AST * q < K) ? ra[q] : T(0);
#pragma unroll
            for (int q = 0; q < BL; ++q) Bs[buf][bk + BST * q][br] = (k0 + bk + BST * q < K) ? rb[q] : T(0);
        }
    };
    gload(0);
    sstore(0, 0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += KC) {
        const bool more = k0 + KC < K;
        if (more) gload(k0 + KC);
#pragma unroll
        for (int ks = 0; ks < KC; ks += 4) {
            const int kr = ks + (lane >> 4), cc = lane & 15;
            T af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) af[i] = As[buf][kr][wm + 16 * i + cc];
#pragma unroll
            for (int j = 0; j < FN; ++j) bf[j] = Bs[buf][kr][wn + 16 * j + cc];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<T>::mma(af[i], bf[j], acc[i][j]);
        }
        if (more) sstore(buf ^ 1, k0 + KC);
        __syncthreads();
        buf ^= 1;
    }
    const T alpha = (t.flags & GT_ALPHA_POS) ? T(1) : T(-1);
    const bool beta0 = (t.flags & GT_BETA0) != 0;
    // waves 0 and 1 keep their quadrant's parts in the A buffers, waves 2 and 3 in the B buffers;
    // FP32 stores its 128x128 tiles in four parts (16 loads of C per lane in flight, which keeps
    // the kernel within 128 registers: four workgroups per CU), the rest in two
    constexpr int NP = (sizeof(T) == 4 && FN == 4) ? 4 : 2, RE = WM * WN / NP;
    static_assert(2 * RE <= 2 * KC * (TM + PAD) && 2 * RE <= 2 * KC * (TN + PAD),
                  "epilogue regions exceed the operand buffers");
    T* region = (wave < 2 ? &As[0][0][0] : &Bs[0][0][0]) + (wave & 1) * RE;
    const int npull = pulls ? (t.flags >> GT_PULL_SHIFT) : 0;
    if (npull > 0)
        apply_pulls<T, FM, FN>(acc, alpha, reinterpret_cast<int*>(region), pulls + t.pad, npull, rel, U, wm, wn,
                               lane);
    gemm_store_lds<T, FM, FN, NP>(acc, C, t.ldc, t.M, t.N, wm, wn, lane, alpha, beta0, region, npull > 0);
}

// Tile GEMM (FP64 or FP32, 128x128 or 64x64 tiles) with its operand chunks loaded straight into
// LDS (buffer_load_dwordx4 ... lds: lane l of a wave writes its 16 bytes at the wave's LDS pointer
// + 16 l), four stages of KC k rows in flight, no register staging.  Same task semantics, fragment
// reads, MFMA order and epilogue arithmetic as gemm_tile<T, TM, TM, 16> (k ascending in steps of
// 4), so the results are the same bits.  One instruction fills 1024 bytes: one 128-double row
// (FP64 128x128 tiles, rows padded by 16 doubles: wave w loads k rows w and w + 4 of each operand)
// or RPI = 2 or 4 unpadded rows (wave w loads rows RPI w .. RPI w + RPI - 1).  FP32 rows hold their
// 4-float groups XOR-swizzled by the row (float m of row r at m ^ 16 (r & 3)), which puts the
// four rows of a fragment read in different banks as the padding does in the register-staged
// kernel.  A row of a chunk beyond K reads an out-of-range offset and lands as zeros.  Rows beyond
// the tile's M (N) read the memory that follows in the operand array (they feed only outputs that
// are never stored); a chunk's resource ends TM elements after its last column's start, which the
// factor arrays' allocation tails keep inside the allocation.  The stages are separate LDS arrays
// and the chunk loop is unrolled by four, so the compiler sees that a stage's loads never alias
// the stage being read and waits only where the code says.
#define MFS_DL_STAGE(n) __shared__ __attribute__((aligned(16))) T As##n[KC][LW], Bs##n[KC][LW];
template <typename T, int TM, bool PULL = false>
__global__ __launch_bounds__(256, 2) void k_gemm_dl(const GemmTask* __restrict__ tasks, T* __restrict__ L,
                                                    T* __restrict__ U, const T* __restrict__ inv, int remap,
                                                    const PullEnt* __restrict__ pulls,
                                                    const int32_t* __restrict__ rel) {
    typedef typename Mfma<T>::acc_t acc_t;
    static_assert(TM == 128 || TM == 64, "tile");
    constexpr int ES = (int)sizeof(T), RPI = 1024 / (TM * ES);              // rows per instruction
    constexpr int KC = (ES == 4 && TM == 64) ? 16 : 8;
    constexpr int IPW = KC / (4 * RPI);                                     // instructions per wave
    constexpr bool PADDED = RPI == 1, SWZ = ES == 4;
    constexpr int LW = PADDED ? TM + 16 : TM, WM = TM / 2, WN = TM / 2, FM = WM / 16, FN = WN / 16;
    constexpr int LPC = 2 * IPW;                                            // loads per wave and chunk
    constexpr int GE = 16 / ES;                                             // elements per lane
    static_assert(IPW >= 1 && IPW * 4 * RPI == KC, "chunk shape");
    MFS_DL_STAGE(0) MFS_DL_STAGE(1) MFS_DL_STAGE(2) MFS_DL_STAGE(3)
    const GemmTask t = tasks[xcd_task(blockIdx.x, gridDim.x, remap)];
    const T* A = L + t.a_off;
    const T* B = (t.flags & GT_B_IN_INV) ? inv + t.b_off : L + t.b_off;
    T* C = (t.flags & GT_C_IN_U) ? U + t.c_off : L + t.c_off;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = (wave & 1) * WM, wn = (wave >> 1) * WN;
    const int lda = t.lda, ldb = t.ldb, K = t.K;
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
    const int nch = (K + KC - 1) / KC;
    auto issue = [&](int c, T (*as)[LW], T (*bs)[LW]) {
        // (a chunk past the last one: every row out of range, an empty resource)
        const int k0 = max(0, min(c, nch - 1)) * KC, kv = c < nch ? min(KC, K - k0) : 0;
        const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(A + (int64_t)k0 * lda), (short)0, kv > 0 ? ((kv - 1) * lda + TM) * ES : 0, kBufWord3);
        const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(B + (int64_t)k0 * ldb), (short)0, kv > 0 ? ((kv - 1) * ldb + TM) * ES : 0, kBufWord3);
        if constexpr (PADDED) {
#pragma unroll
            for (int h = 0; h < IPW; ++h) {
                const int kk = wave + 4 * h;
                const int sa = kk < kv ? kk * lda * ES : 0x7fff0000, sb = kk < kv ? kk * ldb * ES : 0x7fff0000;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rA, (__attribute__((address_space(3))) void*)&as[kk][0], 16, lane * 16, sa, 0, 0);
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rB, (__attribute__((address_space(3))) void*)&bs[kk][0], 16, lane * 16, sb, 0, 0);
            }
        } else {
#pragma unroll
            for (int h = 0; h < IPW; ++h) {
                const int r0 = RPI * (wave * IPW + h);                // first row of the instruction
                const int kk = r0 + lane / (64 / RPI), g = lane % (64 / RPI);
                const int m = (SWZ ? (g ^ (4 * (kk & 3))) : g) * GE;   // first element this lane fetches
                const int va = kk < kv ? (kk * lda + m) * ES : 0x7fff0000;
                const int vb = kk < kv ? (kk * ldb + m) * ES : 0x7fff0000;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rA, (__attribute__((address_space(3))) void*)&as[r0][0], 16, va, 0, 0, 0);
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rB, (__attribute__((address_space(3))) void*)&bs[r0][0], 16, vb, 0, 0, 0);
            }
        }
    };
    auto compute = [&](const T (*as)[LW], const T (*bs)[LW]) {
#pragma unroll
        for (int ks = 0; ks < KC; ks += 4) {
            const int kr = ks + (lane >> 4), cc = lane & 15, sx = SWZ ? 16 * (kr & 3) : 0;
            T af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) af[i] = as[kr][((wm + 16 * i) ^ sx) + cc];
#pragma unroll
            for (int j = 0; j < FN; ++j) bf[j] = bs[kr][((wn + 16 * j) ^ sx) + cc];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<T>::mma(af[i], bf[j], acc[i][j]);
        }
    };
    // chunk c: wait until this wave's loads of c are done (the two later chunks' loads remain:
    // every step issues one chunk, past the end as empty loads, so the count is fixed) and its
    // LDS reads of chunk c - 1 have returned (the scheduler may not move them past this point),
    // barrier (every wave's loads of c have landed; every wave is done reading chunk c - 1), issue
    // chunk c + 3 into the stage of chunk c - 1, compute chunk c
    auto step = [&](int c, T (*as)[LW], T (*bs)[LW], T (*an)[LW], T (*bn)[LW]) {
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(0x0070 | (2 * LPC));   // vmcnt(2 LPC), lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        issue(c + 3, an, bn);
        if (c < nch) compute(as, bs);
    };
    issue(0, As0, Bs0);
    issue(1, As1, Bs1);
    issue(2, As2, Bs2);
    for (int c = 0; c < nch; c += 4) {
        step(c, As0, Bs0, As3, Bs3);
        step(c + 1, As1, Bs1, As0, Bs0);
        step(c + 2, As2, Bs2, As1, Bs1);
        step(c + 3, As3, Bs3, As2, Bs2);
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);   // the empty loads past the end have landed too
    __syncthreads();         // every wave is done with the operand stages
    const T alpha = (t.flags & GT_ALPHA_POS) ? T(1) : T(-1);
    const bool beta0 = (t.flags & GT_BETA0) != 0;
    // epilogue per wave in its own stage array: 128x128 tiles in four 64x16 parts, 64x64 tiles in
    // two 32x16 parts
    constexpr int NP = TM == 128 ? 4 : 2;
    static_assert(WM * WN / NP <= KC * LW, "epilogue region");
    T* region = wave == 0 ? &As0[0][0] : wave == 1 ? &As1[0][0] : wave == 2 ? &Bs0[0][0] : &Bs1[0][0];
    const int npull = PULL ? (t.flags >> GT_PULL_SHIFT) : 0;
    if (npull > 0)
        apply_pulls<T, FM, FN>(acc, alpha, reinterpret_cast<int*>(region), pulls + t.pad, npull, rel, U, wm, wn,
                               lane);
    gemm_store_lds<T, FM, FN, NP>(acc, C, t.ldc, t.M, t.N, wm, wn, lane, alpha, beta0, region, npull > 0);
}
#undef MFS_DL_STAGE

// (second launch bound: at least 2 waves per SIMD, so two 128x128 workgroups share a CU)
template <typename T, int TM, int TN, int KC, bool PULL = false>
__global__ __launch_bounds__(256, 2) void k_gemm(const GemmTask* __restrict__ tasks,
                                                 T* __restrict__ L, T* __restrict__ U,
                                                 const T* __restrict__ inv, int remap,
                                                 const PullEnt* __restrict__ pulls,
                                                 const int32_t* __restrict__ rel) {
    constexpr int PAD = gemm_pad<T, TM>();
    __shared__ T As[2][KC][TM + PAD];
    __shared__ T Bs[2][KC][TN + PAD];
    // (the pull epilogue only in the Schur-update instantiation: its registers would lower the
    // occupancy of every other launch)
    gemm_tile<T, TM, TN, KC>(tasks[xcd_task(blockIdx.x, gridDim.x, remap)], L, U, inv, As, Bs,
                             PULL ? pulls : nullptr, rel);
}

// Fused diagonal tile + TRSM of one step: workgroup = one 64-row tile below the diagonal tile
// (or a diagonal tile with no rows below).  The tile is factored in LDS by every workgroup of its
// group (potrf_tile with the shared-tile counter), then the row tile is solved in place,
// L21 := A21 X^T, with X = L11^-1 straight from LDS: the same MFMA products in the same k order as
// the separate TRSM launch of the FP64 path.  The row tile's MFMA operands are loaded from global
// memory after the factorization (it is this workgroup's alone), so the kernel holds only the tile
// and its inverse in LDS (66.6 KB), and its registers stay within half a SIMD's file: a CU running
// one 128x128 FP64 GEMM workgroup (64 KB, at most 256 registers per wave) of a concurrent launch
// still has room for it.
template <typename T>
__global__ __launch_bounds__(256, 2) void k_panel(const PanelItem* __restrict__ items,
                                               const PotrfTask* __restrict__ ptasks,
                                               const GemmTask* __restrict__ gtasks, T* __restrict__ L,
                                               T* __restrict__ inv, int* __restrict__ status,
                                               int* __restrict__ counters) {
    typedef typename Mfma<T>::acc_t acc_t;
    __shared__ double a[TS][TS + 1];
    __shared__ double x[TS][TS + 1];
    __shared__ int flag;
    const PanelItem it = items[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool rows = it.trsm >= 0;
    GemmTask g{};
    if (rows) g = gtasks[it.trsm];
    potrf_tile<T>(ptasks[it.potrf], L, inv, status, a, x, it.count > 1 ? counters + it.potrf : nullptr,
                  it.count, &flag);
    if (!rows) return;
    T* C = L + g.a_off;         // in place (c_off == a_off)
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32, cc = lane & 15;
    // A21(r, k) for the wave's rows wm + 16 i + cc and the 16 k steps, zero beyond M x K; all loads
    // in one round trip
    T ag[TS / 4][2];
#pragma unroll
    for (int s4 = 0; s4 < TS / 4; ++s4) {
        const int kr = 4 * s4 + (lane >> 4);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int r = wm + 16 * i + cc;
            ag[s4][i] = (r < g.M && kr < g.K) ? C[(int64_t)kr * g.lda + r] : T(0);
        }
    }
    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < TS; ks += 4) {
        const int kr = ks + (lane >> 4);
        T af[2], bf[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) af[i] = ag[ks / 4][i];
#pragma unroll
        for (int j = 0; j < 2; ++j) bf[j] = (T)x[kr][wn + 16 * j + cc];  // B(n, k) = X(n, k), rounded to T
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = Mfma<T>::mma(af[i], bf[j], acc[i][j]);
    }
    gemm_store_acc<T, 2, 2>(acc, C, g.ldc, g.M, g.N, wm, wn, lane, T(1), true);
}

// ------------------------------------------------------------------------------------------
// Fused level kernel: each workgroup factors whole small supernodes (one diagonal tile wide) of
// one elimination-tree level, striding over the level's task list (persistent: the grid is sized
// to the machine, not to the task count; tasks are sorted by cost so the stride balances).  Per
// supernode: the children's update matrices are pulled into the panel (fixed child order), the
// diagonal tile is factored (potrf_tile), the rows below are solved with the tile inverse, the
// update matrix is formed with beta = 0, then the children's columns that land in the update
// matrix are added.  A task writes only its own panel and update matrix, so workgroups never
// touch the same entry; within a task, barriers order the stages.
// ------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void pull_children(const FusedTask& f, bool to_u, const int32_t* __restrict__ children,
                                              const int32_t* __restrict__ sdeg, const int64_t* __restrict__ rel_off,
                                              const int32_t* __restrict__ rel, const int64_t* __restrict__ uoff,
                                              T* __restrict__ U, T* __restrict__ L) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t ldp = f.W + f.D;
    for (int q = f.child_begin; q < f.child_end; ++q) {
        const int c = children[q], dc = sdeg[c];
        const int32_t* relc = rel + rel_off[c];
        const int64_t Dc = 3 * (int64_t)dc;
        const T* uc = U + uoff[c];
        for (int kj = wave; kj < dc; kj += 4) {      // one column block per wave: distinct columns
            const int tj = relc[kj];
            if ((tj >= f.wblk) != to_u) continue;
            for (int ki = kj + lane; ki < dc; ki += 64) {
                const int ti = relc[ki];
                T* dst;
                int64_t ldd;
                if (tj < f.wblk) {
                    dst = L + f.panel_off + (int64_t)(3 * tj) * ldp + 3 * ti;
                    ldd = ldp;
                } else {
                    dst = U + f.u_off + (int64_t)(3 * (tj - f.wblk)) * f.D + 3 * (ti - f.wblk);
                    ldd = f.D;
                }
                const T* src = uc + (3 * (int64_t)kj) * Dc + 3 * (int64_t)ki;
                if (ki == kj) {
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                        for (int r = cc; r < 3; ++r) dst[cc * ldd + r] += src[cc * Dc + r];
                } else {
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                        for (int r = 0; r < 3; ++r) dst[cc * ldd + r] += src[cc * Dc + r];
                }
            }
        }
        __syncthreads();    // the next child may add to the same entries
    }
}

template <typename T>
__global__ __launch_bounds__(256, 2) void k_fused(const FusedTask* __restrict__ tasks, int ntasks,
                                                  const int32_t* __restrict__ children,
                                                  const int32_t* __restrict__ sdeg,
                                                  const int64_t* __restrict__ rel_off,
                                                  const int32_t* __restrict__ rel,
                                                  const int64_t* __restrict__ uoff, T* __restrict__ L,
                                                  T* __restrict__ U, T* __restrict__ inv,
                                                  int* __restrict__ status) {
    constexpr int KC = 16;
    __shared__ __attribute__((aligned(16))) double smem[2 * TS * (TS + 1)];
    double (*a)[TS + 1] = reinterpret_cast<double (*)[TS + 1]>(smem);
    double (*x)[TS + 1] = reinterpret_cast<double (*)[TS + 1]>(smem + TS * (TS + 1));
    constexpr int PAD = gemm_pad<T, TS>();
    static_assert(4 * KC * (TS + PAD) * sizeof(T) <= sizeof(smem), "tile GEMM chunks fit the fused kernel's LDS");
    T* smem_t = reinterpret_cast<T*>(smem);
    T (*As)[KC][TS + PAD] = reinterpret_cast<T (*)[KC][TS + PAD]>(smem_t);
    T (*Bs)[KC][TS + PAD] = reinterpret_cast<T (*)[KC][TS + PAD]>(smem_t + 2 * KC * (TS + PAD));
    for (int item = blockIdx.x; item < ntasks; item += gridDim.x) {
        const FusedTask f = tasks[item];
        const int32_t ld = f.W + f.D;
        pull_children<T>(f, false, children, sdeg, rel_off, rel, uoff, U, L);
        PotrfTask pt;
        pt.diag_off = f.panel_off;
        pt.ld = ld;
        pt.kb = f.W;
        pt.slot = f.slot;
        pt.col0 = f.col0;
        potrf_tile<T>(pt, L, inv, status, a, x);
        __syncthreads();    // tile factor and inverse stored; LDS free for the tile GEMMs
        for (int r0 = f.W; r0 < ld; r0 += TS) {      // TRSM: L21 := A21 X^T, in place
            GemmTask g{};
            g.a_off = f.panel_off + r0;
            g.lda = ld;
            g.b_off = (int64_t)f.slot * TS * TS;
            g.ldb = TS;
            g.c_off = g.a_off;
            g.ldc = ld;
            g.M = min(TS, ld - r0);
            g.N = f.W;
            g.K = f.W;
            g.flags = GT_BETA0 | GT_ALPHA_POS | GT_B_IN_INV;
            gemm_tile<T, TS, TS, KC>(g, L, U, inv, As, Bs);
        }
        __syncthreads();    // L21 complete
        for (int c0 = 0; c0 < f.D; c0 += TS)          // U = -L21 L21^T, lower tiles
            for (int r0 = c0; r0 < f.D; r0 += TS) {
                GemmTask g{};
                g.a_off = f.panel_off + f.W + r0;
                g.lda = ld;
                g.b_off = f.panel_off + f.W + c0;
                g.ldb = ld;
                g.c_off = f.u_off + (int64_t)c0 * f.D + r0;
                g.ldc = f.D;
                g.M = min(TS, f.D - r0);
                g.N = min(TS, f.D - c0);
                g.K = f.W;
                g.flags = GT_C_IN_U | GT_BETA0;
                gemm_tile<T, TS, TS, KC>(g, L, U, inv, As, Bs);
            }
        __syncthreads();    // U written before the children's U columns are added
        pull_children<T>(f, true, children, sdeg, rel_off, rel, uoff, U, L);
    }
}

// executed flops of a task range (2 M N K per tile task)
// Second-generation tile GEMM (same task format and semantics as k_gemm).  Operand tiles sit in
// LDS k-contiguous per row (row stride LDK: 16 k values plus padding that keeps the 16-byte
// accesses of eight consecutive rows in distinct banks).  The MFMA k index of step s is mapped to
// chunk column 4 g + s for lane group g = lane / 16, so each lane fetches its four k values of a
// fragment for the whole chunk with one 16-byte LDS read; A and B use the same mapping, so the
// product is unchanged.  In FP32 the C tile leaves through LDS in two halves for coalesced row
// access; in FP64 the accumulators are stored directly.
// A chunk of KC k values (a multiple of 16) is one global round trip; the MFMA loop walks it in
// blocks of 16 with the fragment mapping above.  Row stride LDK = KC + 16 bytes of padding: an odd
// number of 16-byte units, so 16 consecutive rows start in distinct 16-byte bank groups.
template <typename T, int KC> struct GemmCfg {
    static_assert(KC % 16 == 0, "K chunk");
    static constexpr int LDK = KC + 16 / (int)sizeof(T);
};

template <typename T, int TM, int TN, int KC>
__global__ __launch_bounds__(256, 2) void k_gemm2(const GemmTask* __restrict__ tasks,
                                                  T* __restrict__ L, T* __restrict__ U,
                                                  const T* __restrict__ inv, int remap) {
    typedef typename Mfma<T>::acc_t acc_t;
    typedef T v4 __attribute__((ext_vector_type(4)));
    constexpr int LDK = GemmCfg<T, KC>::LDK;
    constexpr int WM = TM / 2, WN = TN / 2, FM = WM / 16, FN = WN / 16;
    constexpr int AG = TM * (KC / 4) / 256, BG = TN * (KC / 4) / 256;   // (row, 4-k group)s per thread
    constexpr int OPS = 2 * (TM + TN) * LDK, EPI = WN * (TM + 1);
    __shared__ __attribute__((aligned(16))) T smem[OPS > EPI ? OPS : EPI];
    T* const As = smem;                       // [2][TM][LDK]
    T* const Bs = smem + 2 * TM * LDK;        // [2][TN][LDK]
    const GemmTask t = tasks[xcd_task(blockIdx.x, gridDim.x, remap)];
    const T* A = L + t.a_off;
    const T* B = (t.flags & GT_B_IN_INV) ? inv + t.b_off : L + t.b_off;
    T* C = (t.flags & GT_C_IN_U) ? U + t.c_off : L + t.c_off;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * WM, wn = (wave >> 1) * WN;
    const int64_t lda = t.lda, ldb = t.ldb;
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
    // loader: thread -> row tid % TM, k groups tid / TM + (256 / TM) q.  Rows beyond M (N) read
    // row 0 and are zeroed by a select; running pointers advance by KC columns per chunk; only a
    // final partial chunk checks k.
    // loads only fill registers (row 0 for rows beyond M / N, column K - 1 for k beyond K); the
    // masks are applied when the chunk is stored to LDS, after the current chunk's MFMAs
    const int ar = tid % TM, br = tid % TN;
    const int akg = tid / TM, bkg = tid / TN;
    const bool arow = ar < t.M, brow = br < t.N;
    const T* Ac = A + (arow ? ar : 0);
    const T* Bc = B + (brow ? br : 0);
    const T* pa = Ac + (int64_t)(4 * akg) * lda;
    const T* pb = Bc + (int64_t)(4 * bkg) * ldb;
    const int64_t aq = (int64_t)(4 * (256 / TM)) * lda, bq = (int64_t)(4 * (256 / TN)) * ldb;
    v4 ra[AG], rb[BG];
    auto gload = [&](int k0) {
        if (k0 + KC <= t.K) {
#pragma unroll
            for (int q = 0; q < AG; ++q)
#pragma unroll
                for (int u = 0; u < 4; ++u) ra[q][u] = pa[q * aq + u * lda];
#pragma unroll
            for (int q = 0; q < BG; ++q)
#pragma unroll
                for (int u = 0; u < 4; ++u) rb[q][u] = pb[q * bq + u * ldb];
        } else {
#pragma unroll
            for (int q = 0; q < AG; ++q) {
                const int kk = k0 + 4 * (akg + (256 / TM) * q);
#pragma unroll
                for (int u = 0; u < 4; ++u) ra[q][u] = Ac[(int64_t)min(kk + u, t.K - 1) * lda];
            }
#pragma unroll
            for (int q = 0; q < BG; ++q) {
                const int kk = k0 + 4 * (bkg + (256 / TN) * q);
#pragma unroll
                for (int u = 0; u < 4; ++u) rb[q][u] = Bc[(int64_t)min(kk + u, t.K - 1) * ldb];
            }
        }
        pa += KC * lda;
        pb += KC * ldb;
    };
    auto sstore = [&](int buf, int k0) {
#pragma unroll
        for (int q = 0; q < AG; ++q) {
            const int kk = k0 + 4 * (akg + (256 / TM) * q);
            v4 v;
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = (arow && kk + u < t.K) ? ra[q][u] : T(0);
            *(v4*)&As[(buf * TM + ar) * LDK + 4 * (akg + (256 / TM) * q)] = v;
        }
#pragma unroll
        for (int q = 0; q < BG; ++q) {
            const int kk = k0 + 4 * (bkg + (256 / TN) * q);
            v4 v;
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = (brow && kk + u < t.K) ? rb[q][u] : T(0);
            *(v4*)&Bs[(buf * TN + br) * LDK + 4 * (bkg + (256 / TN) * q)] = v;
        }
    };
    gload(0);
    sstore(0, 0);
    __syncthreads();
    int buf = 0;
    const int fr = lane & 15, fg = 4 * (lane >> 4);
    for (int k0 = 0; k0 < t.K; k0 += KC) {
        const bool more = k0 + KC < t.K;
        if (more) gload(k0 + KC);
#pragma unroll
        for (int kb = 0; kb < KC; kb += 16) {
            v4 af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) af[i] = *(const v4*)&As[(buf * TM + wm + 16 * i + fr) * LDK + kb + fg];
#pragma unroll
            for (int j = 0; j < FN; ++j) bf[j] = *(const v4*)&Bs[(buf * TN + wn + 16 * j + fr) * LDK + kb + fg];
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<T>::mma(af[i][s], bf[j][s], acc[i][j]);
        }
        if (more) sstore(buf ^ 1, k0 + KC);
        __syncthreads();
        buf ^= 1;
    }
    const T alpha = (t.flags & GT_ALPHA_POS) ? T(1) : T(-1);
    const bool beta0 = (t.flags & GT_BETA0) != 0;
    const int64_t ldc = t.ldc;
    if constexpr (sizeof(T) == 8) {
        // FP64: straight from the accumulators (measured faster than the LDS round trip)
        gemm_store_acc<T, FM, FN>(acc, C, ldc, t.M, t.N, wm, wn, lane, alpha, beta0);
        return;
    }
    // FP32: column halves [h WN, (h + 1) WN) staged in LDS as Cs[col][row], then coalesced
    T* const Cs = smem;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if ((wave >> 1) == h) {
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        Cs[(16 * j + fr) * (TM + 1) + wm + 16 * i + Mfma<T>::row(lane, q)] = acc[i][j][q];
        }
        __syncthreads();
        // all loads of C first, then all stores: one memory round trip per half
        constexpr int PER = WN * TM / 256;
        T cv[PER];
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int idx = tid + 256 * e, r = idx % TM, col = h * WN + idx / TM;
            cv[e] = (!beta0 && r < t.M && col < t.N) ? C[col * ldc + r] : T(0);
        }
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int idx = tid + 256 * e, r = idx % TM, cl = idx / TM, col = h * WN + cl;
            if (r < t.M && col < t.N) C[col * ldc + r] = cv[e] + alpha * Cs[cl * (TM + 1) + r];
        }
        __syncthreads();
    }
}

static double gemm_work(const Symbolic& S, const TaskRange& r) {
    double w = 0.0;
    for (int64_t q = r.b; q < r.e; ++q) {
        const GemmTask& g = S.gemm_tasks[q];
        w += 2.0 * g.M * g.N * g.K;
    }
    return w;
}
// bytes the tiles of r would move with no reuse between tiles: A (M x K) and B (N x K) panels
// read once per tile, C written (and read first unless beta = 0); elements of `fs` bytes
static double gemm_noreuse_bytes(const Symbolic& S, const TaskRange& r, size_t fs) {
    double b = 0.0;
    for (int64_t q = r.b; q < r.e; ++q) {
        const GemmTask& g = S.gemm_tasks[q];
        b += ((double)(g.M + g.N) * g.K + (double)g.M * g.N * ((g.flags & GT_BETA0) ? 1 : 2)) * (double)fs;
    }
    return b;
}

// tuning knobs: MFS_GEMM_V (1: k_gemm, 2: k_gemm2; default 1) for the
// 128x128 tiles and MFS_GEMM_V64 for the 64x64 tiles (default: as MFS_GEMM_V), MFS_GEMM_KC2 and
// MFS_GEMM_KC64 (k_gemm2's K chunk for the 128x128 and the 64x64 tiles: 16, 32 or 64; default 16),
// MFS_GEMM_REMAP (task-to-XCD mode of xcd_task; default 2 for the FP64 and bfloat16 tiles, 3 for
// the FP32 tiles: measured on the 64^3 step, modes 2 and 3 against 0 FP64 0.2824 -> 0.2802 /
// 0.2809 s, FP32 0.1661 -> 0.1638 / 0.1631 s, BF16x3 0.1181 -> 0.1166 / 0.1169 s; mode 1 is
// slower, FP64 0.2924 and FP32 0.1684)
static int gemm_remap(int element_bytes) {
    static const int v = getenv("MFS_GEMM_REMAP") ? atoi(getenv("MFS_GEMM_REMAP")) : -1;
    if (v >= 0) return v;
    return element_bytes == 4 ? 3 : 2;
}
// (read at every launch, so a process can compare the kernels)
static int gemm_version() {
    const char* v = getenv("MFS_GEMM_V");
    return v ? atoi(v) : 0;
}
static int env_int(const char* name, int dflt) {
    const char* v = getenv(name);
    return v ? atoi(v) : dflt;
}
static int gemm_version64() { return env_int("MFS_GEMM_V64", 0); }
static int gemm_kc2() {
    static const int v = env_int("MFS_GEMM_KC2", 16);
    return v;
}
static int gemm_kc64() {
    static const int v = env_int("MFS_GEMM_KC64", 16);
    return v;
}
template <typename T, int TM>
static void launch_gemm2(int kc, unsigned grid, const GemmTask* tasks, T* L, T* U, const T* inv, int rm,
                         hipStream_t s) {
    // larger chunks only where two workgroups still share a CU's LDS (measured: no gain on the
    // 64^3 schedule; kept as a tuning knob)
    constexpr bool k32_fits = 2 * 2 * (2 * TM) * (32 + 16 / (int)sizeof(T)) * (int)sizeof(T) <= 160 * 1024;
    constexpr bool k64_fits = 2 * 2 * (2 * TM) * (64 + 16 / (int)sizeof(T)) * (int)sizeof(T) <= 160 * 1024;
    if constexpr (k64_fits) {
        if (kc >= 64) {
            k_gemm2<T, TM, TM, 64><<<grid, 256, 0, s>>>(tasks, L, U, inv, rm);
            LAUNCH_CHECK();
            return;
        }
    }
    if constexpr (k32_fits) {
        if (kc >= 32) {
            k_gemm2<T, TM, TM, 32><<<grid, 256, 0, s>>>(tasks, L, U, inv, rm);
            LAUNCH_CHECK();
            return;
        }
    }
    k_gemm2<T, TM, TM, 16><<<grid, 256, 0, s>>>(tasks, L, U, inv, rm);
    LAUNCH_CHECK();
}
template <typename T>
static void launch_gemm(const TaskRange& r, const GemmTask* tasks, T* L, T* U, const T* inv, hipStream_t s,
                        const PullEnt* pulls = nullptr, const int32_t* rel = nullptr) {
    const int rm = gemm_remap((int)sizeof(T));
    // default: the 128x128 tiles load their operands straight into LDS (k_gemm_dl, version 3), and
    // so do the FP64 64x64 tiles; the FP32 64x64 tiles stage them through registers (the k-major
    // buffer-load kernel, version 1: eight workgroups per CU against five; measured FP32 step
    // 156.1 -> 154.0 ms with the direct 128x128 tiles, three runs each).  The k-contiguous kernel,
    // version 2, remains selectable.
    const int gv = gemm_version(), gv64 = gemm_version64();
    const int ver = gv ? gv : 3;
    const int ver64 = gv64 ? gv64 : (gv ? gv : (sizeof(T) == 8 ? 3 : 1));
    if (r.m > r.b) {
        if (ver == 2 && !pulls) launch_gemm2<T, 128>(gemm_kc2(), (unsigned)(r.m - r.b), tasks + r.b, L, U, inv, rm, s);
        else if (ver == 3) {
            if (pulls) k_gemm_dl<T, 128, true><<<(unsigned)(r.m - r.b), 256, 0, s>>>(tasks + r.b, L, U, inv, rm, pulls, rel);
            else k_gemm_dl<T, 128><<<(unsigned)(r.m - r.b), 256, 0, s>>>(tasks + r.b, L, U, inv, rm, nullptr, nullptr);
            LAUNCH_CHECK();
        } else {
            if (pulls) k_gemm<T, 128, 128, 16, true><<<(unsigned)(r.m - r.b), 256, 0, s>>>(tasks + r.b, L, U, inv, rm, pulls, rel);
            else k_gemm<T, 128, 128, 16><<<(unsigned)(r.m - r.b), 256, 0, s>>>(tasks + r.b, L, U, inv, rm, nullptr, nullptr);
            LAUNCH_CHECK();
        }
    }
    if (r.e > r.m) {
        if (ver64 == 2 && !pulls) launch_gemm2<T, 64>(gemm_kc64(), (unsigned)(r.e - r.m), tasks + r.m, L, U, inv, rm, s);
        else if (ver64 == 3) {
            if (pulls) k_gemm_dl<T, 64, true><<<(unsigned)(r.e - r.m), 256, 0, s>>>(tasks + r.m, L, U, inv, rm, pulls, rel);
            else k_gemm_dl<T, 64><<<(unsigned)(r.e - r.m), 256, 0, s>>>(tasks + r.m, L, U, inv, rm, nullptr, nullptr);
            LAUNCH_CHECK();
        } else {
            if (pulls) k_gemm<T, 64, 64, 16, true><<<(unsigned)(r.e - r.m), 256, 0, s>>>(tasks + r.m, L, U, inv, rm, pulls, rel);
            else k_gemm<T, 64, 64, 16><<<(unsigned)(r.e - r.m), 256, 0, s>>>(tasks + r.m, L, U, inv, rm, nullptr, nullptr);
            LAUNCH_CHECK();
        }
    }
}

// Tile GEMM for the BF16 Schur-update variant of the FP32 factorization: same task format and
// semantics as k_gemm on FP32 panels, but the operands are rounded to bfloat16 while they are
// staged into LDS and multiplied by v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  SPLIT
// keeps the rounding residual as a second bfloat16 term, v = hi + lo, and accumulates
// hi*hi + hi*lo + lo*hi (three MFMAs), which carries about 16 significant bits of each operand
// instead of 8.  LDS holds each operand tile k-contiguous per row (32 k plus 8 of padding, so
// the 16-byte fragment reads of eight consecutive rows fall in distinct banks); lane l reads
// row l % 16 and the eight k values 8 (l / 16) .. 8 (l / 16) + 7, the same k slots for A and B.
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// NT threads (2 x NT/128 waves): 256 for 128x128 and 64x64 tiles, 512 for 256x256 tiles (half the
// operand loads and conversions per flop; its two LDS buffers fill a CU's 160 KB).
template <int TM, int TN, bool SPLIT, int NT = 256, bool PULL = false>
__global__ __launch_bounds__(NT, 512 / NT) void k_gemm_bf(const GemmTask* __restrict__ tasks,
                                                          float* __restrict__ L, float* __restrict__ U,
                                                          const float* __restrict__ inv, int remap,
                                                          const PullEnt* __restrict__ pulls,
                                                          const int32_t* __restrict__ rel) {
    typedef f4 acc_t;
    constexpr int KC = 32, LDK = 40, NP = SPLIT ? 2 : 1;
    constexpr int WGY = NT / 128;                          // wave columns
    constexpr int WM = TM / 2, WN = TN / WGY, FM = WM / 16, FN = WN / 16;
    constexpr int AKT = NT / TM, BKT = NT / TN;            // threads per operand row
    constexpr int AKE = KC / AKT, BKE = KC / BKT;          // k values per thread and chunk
    constexpr int STAGE = NP * (TM + TN) * LDK;            // bf16 per buffer
    __shared__ __attribute__((aligned(16))) __bf16 smem[2 * STAGE];
    const GemmTask t = tasks[xcd_task(blockIdx.x, gridDim.x, remap)];
    const float* A = L + t.a_off;
    const float* B = (t.flags & GT_B_IN_INV) ? inv + t.b_off : L + t.b_off;
    float* C = (t.flags & GT_C_IN_U) ? U + t.c_off : L + t.c_off;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * WM, wn = (wave >> 1) * WN;
    const int lda = t.lda, ldb = t.ldb, K = t.K;
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
    const int ar = tid % TM, br = tid % TN, ak0 = (tid / TM) * AKE, bk0 = (tid / TN) * BKE;
    // Raw buffer loads as in gemm_tile: per chunk a resource based at column k0, a per-thread row
    // offset and uniform per-load column offsets.  Rows beyond M / N read row 0 (they only feed
    // outputs that are never stored); the partial chunk clamps k to K - 1 and zeroes k >= K.
    const int arow = ar < t.M ? ar : 0, brow = br < t.N ? br : 0;
    const int va = (arow + ak0 * lda) * 4, vb = (brow + bk0 * ldb) * 4;
    const int ea = KC * lda * 4, eb = KC * ldb * 4;
    float ra[AKE], rb[BKE];
    auto gload = [&](int k0) {
        const __amdgpu_buffer_rsrc_t rA =
            __builtin_amdgcn_make_buffer_rsrc((void*)(A + (int64_t)k0 * lda), (short)0, ea, kBufWord3);
        const __amdgpu_buffer_rsrc_t rB =
            __builtin_amdgcn_make_buffer_rsrc((void*)(B + (int64_t)k0 * ldb), (short)0, eb, kBufWord3);
        if (k0 + KC <= K) {
#pragma unroll
            for (int u = 0; u < AKE; ++u) ra[u] = buf_load<float>(rA, va, u * lda * 4);
#pragma unroll
            for (int u = 0; u < BKE; ++u) rb[u] = buf_load<float>(rB, vb, u * ldb * 4);
        } else {
            const int kl = K - 1 - k0;
#pragma unroll
            for (int u = 0; u < AKE; ++u) ra[u] = buf_load<float>(rA, (arow + min(ak0 + u, kl) * lda) * 4, 0);
#pragma unroll
            for (int u = 0; u < BKE; ++u) rb[u] = buf_load<float>(rB, (brow + min(bk0 + u, kl) * ldb) * 4, 0);
        }
    };
    auto put = [&](__bf16* base, int row, int k0_, const float* v, int n) {
#pragma unroll
        for (int g = 0; g < n / 8; ++g) {
            bf8 hi;
#pragma unroll
            for (int u = 0; u < 8; ++u) hi[u] = (__bf16)v[8 * g + u];
            *(bf8*)&base[row * LDK + k0_ + 8 * g] = hi;
            if constexpr (SPLIT) {
                bf8 lo;
#pragma unroll
                for (int u = 0; u < 8; ++u) lo[u] = (__bf16)(v[8 * g + u] - (float)hi[u]);
                *(bf8*)&base[(TM + TN) * LDK + row * LDK + k0_ + 8 * g] = lo;
            }
        }
    };
    auto sstore = [&](int buf, int k0) {
        __bf16* st = smem + buf * STAGE;
        if (k0 + KC <= K) {
            put(st, ar, ak0, ra, AKE);
            put(st + TM * LDK, br, bk0, rb, BKE);
        } else {
            float ma[AKE], mb[BKE];
#pragma unroll
            for (int u = 0; u < AKE; ++u) ma[u] = (k0 + ak0 + u < K) ? ra[u] : 0.0f;
#pragma unroll
            for (int u = 0; u < BKE; ++u) mb[u] = (k0 + bk0 + u < K) ? rb[u] : 0.0f;
            put(st, ar, ak0, ma, AKE);
            put(st + TM * LDK, br, bk0, mb, BKE);
        }
    };
    gload(0);
    sstore(0, 0);
    __syncthreads();
    int buf = 0;
    const int fr = lane & 15, fk = 8 * (lane >> 4);
    for (int k0 = 0; k0 < K; k0 += KC) {
        const bool more = k0 + KC < K;
        if (more) gload(k0 + KC);
        const __bf16* st = smem + buf * STAGE;
        const __bf16* sl = st + (TM + TN) * LDK;
        // B fragments once per chunk, A fragments one row block at a time (registers); every
        // accumulator still receives lo*hi, hi*lo, then hi*hi in each chunk
        bf8 bh[FN], bl[SPLIT ? FN : 1];
#pragma unroll
        for (int j = 0; j < FN; ++j) bh[j] = *(const bf8*)&st[(TM + wn + 16 * j + fr) * LDK + fk];
        if constexpr (SPLIT) {
#pragma unroll
            for (int j = 0; j < FN; ++j) bl[j] = *(const bf8*)&sl[(TM + wn + 16 * j + fr) * LDK + fk];
        }
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const bf8 ah = *(const bf8*)&st[(wm + 16 * i + fr) * LDK + fk];
            if constexpr (SPLIT) {
                const bf8 al = *(const bf8*)&sl[(wm + 16 * i + fr) * LDK + fk];
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], acc[i][j], 0, 0, 0);
                }
            }
#pragma unroll
            for (int j = 0; j < FN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], acc[i][j], 0, 0, 0);
        }
        if (more) sstore(buf ^ 1, k0 + KC);
        __syncthreads();
        buf ^= 1;
    }
    const float alpha = (t.flags & GT_ALPHA_POS) ? 1.0f : -1.0f;
    // C through LDS one 16-column fragment per part (each wave's region in the operand buffers)
    constexpr int RE = WM * 16;
    static_assert((NT / 64) * RE * 4 <= 2 * STAGE * 2, "epilogue regions exceed the operand buffers");
    float* region = reinterpret_cast<float*>(smem) + wave * RE;
    // a Schur-update tile's children lists (the Schur-update instantiation only)
    const int npull = PULL ? (t.flags >> GT_PULL_SHIFT) : 0;
    if (npull > 0)
        apply_pulls<float, FM, FN>(acc, alpha, reinterpret_cast<int*>(region), pulls + t.pad, npull, rel, U, wm,
                                   wn, lane);
    gemm_store_lds<float, FM, FN, FN>(acc, C, t.ldc, t.M, t.N, wm, wn, lane, alpha, (t.flags & GT_BETA0) != 0,
                                      region, npull > 0);
}

// big256: the range's large tiles are 256x256 (Schur ranges of an analysis with schur_tile 256)
template <int TM, int TN, int NT = 256>
static void launch_gemm_bf1(bool split, unsigned grid, const GemmTask* tasks, float* L, float* U, const float* inv,
                            int rm, const PullEnt* pulls, const int32_t* rel, hipStream_t s) {
    if (pulls) {
        if (split) k_gemm_bf<TM, TN, true, NT, true><<<grid, NT, 0, s>>>(tasks, L, U, inv, rm, pulls, rel);
        else k_gemm_bf<TM, TN, false, NT, true><<<grid, NT, 0, s>>>(tasks, L, U, inv, rm, pulls, rel);
    } else {
        if (split) k_gemm_bf<TM, TN, true, NT><<<grid, NT, 0, s>>>(tasks, L, U, inv, rm, nullptr, nullptr);
        else k_gemm_bf<TM, TN, false, NT><<<grid, NT, 0, s>>>(tasks, L, U, inv, rm, nullptr, nullptr);
    }
    LAUNCH_CHECK();
}
static void launch_gemm_bf(const TaskRange& r, const GemmTask* tasks, float* L, float* U, const float* inv,
                           bool split, bool big256, hipStream_t s, const PullEnt* pulls = nullptr,
                           const int32_t* rel = nullptr) {
    const int rm = gemm_remap(2);
    if (r.m > r.b && big256)
        launch_gemm_bf1<256, 256, 512>(split, (unsigned)(r.m - r.b), tasks + r.b, L, U, inv, rm, pulls, rel, s);
    else if (r.m > r.b)
        launch_gemm_bf1<128, 128>(split, (unsigned)(r.m - r.b), tasks + r.b, L, U, inv, rm, pulls, rel, s);
    if (r.e > r.m) launch_gemm_bf1<64, 64>(split, (unsigned)(r.e - r.m), tasks + r.m, L, U, inv, rm, pulls, rel, s);
}

// ------------------------------------------------------------------------------------------
// Solve kernels (x = factor-order reduced vector of the factor type; one RHS per blockIdx.y)
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_fgather(const SolveGatherTask* __restrict__ tasks,
                                                 const int32_t* __restrict__ children,
                                                 const int32_t* __restrict__ sdeg,
                                                 const int64_t* __restrict__ rel_off,
                                                 const int32_t* __restrict__ rel,
                                                 const int64_t* __restrict__ usoff,
                                                 T* __restrict__ xall, int64_t ldx,
                                                 T* __restrict__ uall, int64_t ldu) {
    const SolveGatherTask t = tasks[blockIdx.x];
    T* x = xall + blockIdx.y * ldx;
    T* u = uall + blockIdx.y * ldu;
    T* us = u + t.u_off;
    for (int i = threadIdx.x; i < t.d; i += blockDim.x) us[i] = T(0);
    __syncthreads();
    for (int q = t.child_begin; q < t.child_end; ++q) {
        const int c = children[q];
        const int Dc = 3 * sdeg[c];
        const T* uc = u + usoff[c];
        const int32_t* relc = rel + rel_off[c];
        // 8 entries per thread and pass: all index and value loads, then all target reads, then
        // the stores (a child's targets are distinct, so the batch has no conflicts; children
        // stay in order, one per barrier)
        constexpr int UB = 8;
        for (int k0 = threadIdx.x; k0 < Dc; k0 += UB * blockDim.x) {
            int tg[UB];
            T v[UB], cur[UB];
#pragma unroll
            for (int e = 0; e < UB; ++e) {
                const int k = min(k0 + e * (int)blockDim.x, Dc - 1);
                tg[e] = 3 * relc[k / 3] + k % 3;
                v[e] = uc[k];
            }
#pragma unroll
            for (int e = 0; e < UB; ++e)
                if (k0 + e * (int)blockDim.x < Dc) cur[e] = tg[e] < t.w ? x[t.x_off + tg[e]] : us[tg[e] - t.w];
#pragma unroll
            for (int e = 0; e < UB; ++e)
                if (k0 + e * (int)blockDim.x < Dc) {
                    if (tg[e] < t.w) x[t.x_off + tg[e]] = cur[e] - v[e];
                    else us[tg[e] - t.w] = cur[e] + v[e];
                }
        }
        __syncthreads();
    }
}

// Backward gather: u = x at the supernode's structure rows (final values of its ancestors)
template <typename T>
__global__ __launch_bounds__(256) void k_bgather(const SolveGatherTask* __restrict__ tasks,
                                                 const int32_t* __restrict__ rows,
                                                 const T* __restrict__ xall, int64_t ldx,
                                                 T* __restrict__ uall, int64_t ldu) {
    const SolveGatherTask t = tasks[blockIdx.x];
    const T* x = xall + blockIdx.y * ldx;
    T* u = uall + blockIdx.y * ldu + t.u_off;
    const int32_t* rw = rows + t.rows_off;
    // 8 entries per thread and pass (index loads, then value loads, then stores)
    constexpr int UB = 8;
    for (int i0 = threadIdx.x; i0 < t.d; i0 += UB * blockDim.x) {
        int64_t src[UB];
        T v[UB];
#pragma unroll
        for (int e = 0; e < UB; ++e) {
            const int i = min(i0 + e * (int)blockDim.x, t.d - 1);
            src[e] = 3 * (int64_t)rw[i / 3] + i % 3;
        }
#pragma unroll
        for (int e = 0; e < UB; ++e) v[e] = x[src[e]];
#pragma unroll
        for (int e = 0; e < UB; ++e)
            if (i0 + e * (int)blockDim.x < t.d) u[i0 + e * blockDim.x] = v[e];
    }
}

// deterministic wave-wide sum (fixed butterfly); every lane ends with the same value
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block inverses X = L_bb^{-1} (lower triangular, nb x nb, column-major, ld nb), one workgroup per
// 64-column tile j: X_jj = stored tile inverse, then for i > j
//   X_ij = -X_ii * sum_{m=j}^{i-1} L_im X_mj
// with 64x64x64 tile products on MFMA (2x2 waves, each a 32x32 quadrant of 16x16 fragments).
// The tiles X_mj this workgroup wrote are read back after a barrier (same workgroup, same CU).
template <typename T>
__global__ __launch_bounds__(256) void k_blkinv(const SolveTile* __restrict__ tiles,
                                                const SolveBlock* __restrict__ blocks,
                                                const int32_t* __restrict__ tile_slot,
                                                const T* __restrict__ L, const T* __restrict__ inv,
                                                T* __restrict__ Xb) {
    typedef typename Mfma<T>::acc_t acc_t;
    const SolveTile st = tiles[blockIdx.x];
    const SolveBlock b = blocks[st.blk];
    const int j = st.t, nb = b.nb, nt = (nb + TS - 1) / TS;
    const int64_t ld = b.ld;
    T* X = Xb + b.xb_off;
    const T* Lb = L + b.l_off;
    __shared__ T As[TS][TS + 1];   // As[k][r] = A(r, k)
    __shared__ T Bs[TS][TS + 1];   // Bs[k][c] = B(k, c)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = 32 * (wave & 1), wc = 32 * (wave >> 1);
    const int cj = min(TS, nb - TS * j);
    {
        const T* Vj = inv + (int64_t)tile_slot[b.ts_off + j] * TS * TS;
        for (int idx = tid; idx < TS * TS; idx += 256) {
            const int r = idx & 63, c = idx >> 6;
            if (r < cj && c < cj) X[(int64_t)(TS * j + c) * nb + TS * j + r] = (r >= c) ? Vj[idx] : T(0);
        }
    }
    auto mma_tile = [&](acc_t (&acc)[2][2]) {
#pragma unroll 4
        for (int k = 0; k < TS; k += 4) {
            T a[2], bb[2];
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                a[f] = As[k + (lane >> 4)][wr + 16 * f + (lane & 15)];
                bb[f] = Bs[k + (lane >> 4)][wc + 16 * f + (lane & 15)];
            }
#pragma unroll
            for (int fi = 0; fi < 2; ++fi)
#pragma unroll
                for (int fj = 0; fj < 2; ++fj) acc[fi][fj] = Mfma<T>::mma(a[fi], bb[fj], acc[fi][fj]);
        }
    };
    for (int i = j + 1; i < nt; ++i) {
        const int ri = min(TS, nb - TS * i);
        acc_t acc[2][2];
#pragma unroll
        for (int fi = 0; fi < 2; ++fi)
#pragma unroll
            for (int fj = 0; fj < 2; ++fj)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[fi][fj][q] = T(0);
        // A = L_im, B = X_mj (X_mj of m < i were stored by this workgroup before).  The tiles
        // of step m + 1 are loaded into registers (indices clamped in bounds) while step m's
        // MFMAs run, and masked when stored to LDS.
        // thread: row r = lane of the tile, columns k = wave + 4 q
        constexpr int PQ = TS * TS / 256;
        T ra[PQ], rb[PQ];
        const T* arow_p = Lb + TS * i + min(lane, ri - 1);
        auto gl = [&](int m) {
            const int cm = min(TS, nb - TS * m);
            const T* a = arow_p + (int64_t)(TS * m) * ld;
            const T* bcol = X + (int64_t)(TS * j) * nb + TS * m + min(lane, cm - 1);
#pragma unroll
            for (int q = 0; q < PQ; ++q) {
                const int k = wave + 4 * q;
                ra[q] = a[(int64_t)min(k, cm - 1) * ld];
                rb[q] = bcol[(int64_t)min(k, cj - 1) * nb];
            }
        };
        gl(j);
        for (int m = j; m < i; ++m) {
            const int cm = min(TS, nb - TS * m);
            __syncthreads();
#pragma unroll
            for (int q = 0; q < PQ; ++q) {
                const int k = wave + 4 * q;
                As[k][lane] = (lane < ri && k < cm) ? ra[q] : T(0);      // A(r, k) = L_im
                Bs[lane][k] = (lane < cm && k < cj) ? rb[q] : T(0);      // B(kk = r, c = k) = X_mj
            }
            __syncthreads();
            if (m + 1 < i) gl(m + 1);
            mma_tile(acc);
        }
        __syncthreads();
        // B = the sum (rows k), A = X_ii (lower triangle of the stored tile inverse)
#pragma unroll
        for (int fi = 0; fi < 2; ++fi)
#pragma unroll
            for (int fj = 0; fj < 2; ++fj)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    Bs[wr + 16 * fi + Mfma<T>::row(lane, q)][wc + 16 * fj + (lane & 15)] = acc[fi][fj][q];
        {
            const T* Vi = inv + (int64_t)tile_slot[b.ts_off + i] * TS * TS;
            for (int idx = tid; idx < TS * TS; idx += 256) {
                const int r = idx & 63, k = idx >> 6;
                As[k][r] = (r >= k) ? Vi[idx] : T(0);
            }
        }
        __syncthreads();
#pragma unroll
        for (int fi = 0; fi < 2; ++fi)
#pragma unroll
            for (int fj = 0; fj < 2; ++fj)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[fi][fj][q] = T(0);
        mma_tile(acc);
#pragma unroll
        for (int fi = 0; fi < 2; ++fi)
#pragma unroll
            for (int fj = 0; fj < 2; ++fj)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = wr + 16 * fi + Mfma<T>::row(lane, q), c = wc + 16 * fj + (lane & 15);
                    if (r < ri && c < cj) X[(int64_t)(TS * j + c) * nb + TS * i + r] = -acc[fi][fj][q];
                }
    }
}

// Columns c, c + NW, ..., below nc of a lane's row (ld apart), weighted by yv, added to acc in
// column order: one memory round trip for up to Q columns (past nc the loads repeat column nc - 1
// and are skipped).
template <typename T, int Q, int NW>
__device__ __forceinline__ T gemv_tail(const T* __restrict__ P, int64_t ld, const T* __restrict__ yv, int c,
                                       int nc, T acc) {
    T v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] = P[(int64_t)min(c + NW * q, nc - 1) * ld];
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (c + NW * q < nc) acc += v[q] * yv[c + NW * q];
    return acc;
}

// A lane's row of a column-major panel P (ld) times yv over the columns wave, wave + NW, ... < nc,
// summed in column order: batches of 32 independent loads per memory round trip.
template <typename T, int NW>
__device__ __forceinline__ T gemv_row(const T* __restrict__ P, int64_t ld, const T* __restrict__ yv, int nc,
                                      int wave) {
    T acc = T(0);
    int c = wave;
    for (; c + 31 * NW < nc; c += 32 * NW) {
        T v[32];
#pragma unroll
        for (int q = 0; q < 32; ++q) v[q] = P[(int64_t)(c + NW * q) * ld];
#pragma unroll
        for (int q = 0; q < 32; ++q) acc += v[q] * yv[c + NW * q];
    }
    if (c + 8 * NW < nc) acc = gemv_tail<T, 32, NW>(P, ld, yv, c, nc, acc);
    else if (c < nc) acc = gemv_tail<T, 8, NW>(P, ld, yv, c, nc, acc);
    return acc;
}

// Sum of the NW per-wave partials of row r in wave order (fixed).
template <typename T, int NW>
__device__ __forceinline__ T wave_partials(const T (*red)[TS], int r) {
    T v = red[0][r];
#pragma unroll
    for (int w = 1; w < NW; ++w) v += red[w][r];
    return v;
}

// Forward diagonal: y(rows of tile t) = X(rows, 0 .. 64(t+1)) x_b; lanes over rows, the four waves
// over interleaved columns, partial sums combined in a fixed order.
template <typename T>
__global__ __launch_bounds__(256) void k_sfdiag(const SolveTile* __restrict__ tiles,
                                                    const SolveBlock* __restrict__ blocks,
                                                    const T* __restrict__ Xb,
                                                    const T* __restrict__ xall, T* __restrict__ yall,
                                                    int64_t ldx) {
    constexpr int NW = 4;
    const SolveTile st = tiles[blockIdx.x];
    const SolveBlock b = blocks[st.blk];
    const T* x = xall + blockIdx.y * ldx + b.x_off;
    T* y = yall + blockIdx.y * ldx + b.x_off;
    __shared__ T z[1024];
    __shared__ T red[NW][TS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = TS * st.t, nr = min(TS, b.nb - r0), nc = min(b.nb, r0 + TS);
    for (int c = tid; c < nc; c += 64 * NW) z[c] = x[c];
    __syncthreads();
    red[wave][lane] = lane < nr ? gemv_row<T, NW>(Xb + b.xb_off + r0 + lane, b.nb, z, nc, wave) : T(0);
    __syncthreads();
    if (tid < nr) y[r0 + tid] = wave_partials<T, NW>(red, tid);
}

// Forward update: rows [r0, r0 + nr) below block b: x -= L(rows, b) y_b (rows of the supernode)
// or u += L(rows, b) y_b (structure rows, the update vector passed to the parent).
template <typename T>
__global__ __launch_bounds__(256) void k_sfupd(const SolveRowTask* __restrict__ tasks, int64_t ntasks,
                                                   const T* __restrict__ L,
                                                   const T* __restrict__ yall, T* __restrict__ xall,
                                                   int64_t ldx, T* __restrict__ uall, int64_t ldu) {
    constexpr int NW = 4;
    const int64_t ti = xcd_chunk(blockIdx.x, gridDim.x);
    if (ti >= ntasks) return;                  // the whole workgroup leaves together
    const SolveRowTask t = tasks[ti];
    __shared__ T yv[1024];
    __shared__ T red[NW][TS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const T* y = yall + blockIdx.y * ldx + t.v_off;
    for (int c = tid; c < t.nc; c += 64 * NW) yv[c] = y[c];
    __syncthreads();
    red[wave][lane] = lane < t.nr ? gemv_row<T, NW>(L + t.l_off + lane, t.ld, yv, t.nc, wave) : T(0);
    __syncthreads();
    if (tid < t.nr) {
        const T v = wave_partials<T, NW>(red, tid);
        if (t.mode == 0) xall[blockIdx.y * ldx + t.d_off + tid] -= v;
        else uall[blockIdx.y * ldu + t.d_off + tid] += v;
    }
}

// Quarter-tile forms for launches of few tiles: the 64-row tile (or, backward, 64-column tile)
// of a task is split over 4 workgroups of 16 rows (columns), so a step's bytes are read by 4 times
// as many CUs.  Forward: a lane works on row lane % 16 of its quarter and column stream
// 4 wave + lane / 16 (columns stream, stream + 64, ...); the four streams of a wave are combined
// by shuffles and the waves through LDS, both in a fixed order.
template <typename T>
__device__ __forceinline__ T gemv_row_q(const T* __restrict__ P, int64_t ld, const T* __restrict__ yv, int nc,
                                        int sid) {
    // nc <= 1024: at most 16 columns per stream
    if (sid + 8 * 64 < nc) return gemv_tail<T, 16, 64>(P, ld, yv, sid, nc, T(0));
    if (sid < nc) return gemv_tail<T, 8, 64>(P, ld, yv, sid, nc, T(0));
    return T(0);
}

// Row sums of a 16-row quarter: lanes r, r + 16, r + 32, r + 48 of each wave, then the 16 waves.
template <typename T>
__device__ __forceinline__ bool quarter_sum(T acc, T (*red)[16], T& out) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const T a0 = __shfl(acc, lane & 15, 64), a1 = __shfl(acc, (lane & 15) + 16, 64);
    const T a2 = __shfl(acc, (lane & 15) + 32, 64), a3 = __shfl(acc, (lane & 15) + 48, 64);
    if (lane < 16) red[wave][lane] = ((a0 + a1) + a2) + a3;
    __syncthreads();
    if (tid >= 16) return false;
    T v = red[0][tid];
#pragma unroll
    for (int w = 1; w < 16; ++w) v += red[w][tid];
    out = v;
    return true;
}

template <typename T>
__global__ __launch_bounds__(1024) void k_sfdiag_q(const SolveTile* __restrict__ tiles,
                                                   const SolveBlock* __restrict__ blocks,
                                                   const T* __restrict__ Xb,
                                                   const T* __restrict__ xall, T* __restrict__ yall,
                                                   int64_t ldx) {
    const SolveTile st = tiles[blockIdx.x >> 2];
    const SolveBlock b = blocks[st.blk];
    const T* x = xall + blockIdx.y * ldx + b.x_off;
    T* y = yall + blockIdx.y * ldx + b.x_off;
    __shared__ T z[1024];
    __shared__ T red[16][16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = TS * st.t + 16 * (blockIdx.x & 3), nr = min(16, b.nb - r0), nc = min(b.nb, TS * (st.t + 1));
    if (nr <= 0) return;                       // the whole workgroup leaves together
    for (int c = tid; c < nc; c += 1024) z[c] = x[c];
    __syncthreads();
    const int r = lane & 15;
    const T acc = r < nr ? gemv_row_q<T>(Xb + b.xb_off + r0 + r, b.nb, z, nc, 4 * wave + (lane >> 4)) : T(0);
    T v;
    if (quarter_sum<T>(acc, red, v) && tid < nr) y[r0 + tid] = v;
}

template <typename T>
__global__ __launch_bounds__(1024) void k_sfupd_q(const SolveRowTask* __restrict__ tasks, int64_t ntasks,
                                                  const T* __restrict__ L,
                                                  const T* __restrict__ yall, T* __restrict__ xall,
                                                  int64_t ldx, T* __restrict__ uall, int64_t ldu) {
    const int64_t qi = xcd_chunk(blockIdx.x, gridDim.x);
    if (qi >= 4 * ntasks) return;              // the whole workgroup leaves together
    const SolveRowTask t = tasks[qi >> 2];
    const int q0 = 16 * (int)(qi & 3), nr = min(16, t.nr - q0);
    if (nr <= 0) return;                       // the whole workgroup leaves together
    __shared__ T yv[1024];
    __shared__ T red[16][16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const T* y = yall + blockIdx.y * ldx + t.v_off;
    for (int c = tid; c < t.nc; c += 1024) yv[c] = y[c];
    __syncthreads();
    const int r = lane & 15;
    const T acc = r < nr ? gemv_row_q<T>(L + t.l_off + q0 + r, t.ld, yv, t.nc, 4 * wave + (lane >> 4)) : T(0);
    T v;
    if (quarter_sum<T>(acc, red, v) && tid < nr) {
        if (t.mode == 0) xall[blockIdx.y * ldx + t.d_off + q0 + tid] -= v;
        else uall[blockIdx.y * ldu + t.d_off + q0 + tid] += v;
    }
}

constexpr int SOLVE_RC = 1024;   // maximum rows per backward partial task
// Rows [64 M0, 64 (M0 + MR)) of a task: a wave's columns (at most 4 of the task's <= 16) are
// loaded in one round trip and their products added to acc[j] in row order; rows beyond nr read
// row 0 of the column (in bounds) and are zeroed by the product with x = 0.
template <typename T, int MR>
__device__ __forceinline__ void sbpart_rows(const SolveRowTask& t, const T* __restrict__ L,
                                            const T* __restrict__ src, T (&acc)[4], int M0) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T xv[MR];
#pragma unroll
    for (int m = 0; m < MR; ++m) {
        const int r = lane + 64 * (M0 + m);
        xv[m] = r < t.nr ? src[r] : T(0);
    }
    T v[4][MR];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const T* P = L + t.l_off + (int64_t)min(wave + 4 * j, t.nc - 1) * t.ld;
#pragma unroll
        for (int m = 0; m < MR; ++m) {
            const int r = lane + 64 * (M0 + m);
            v[j][m] = P[r < t.nr ? r : 0];
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int m = 0; m < MR; ++m) acc[j] += v[j][m] * xv[m];
}

// MR chunks of 64 rows (16 as two halves of 8, which keeps the kernel's registers, and so its
// occupancy, at the 8-chunk size)
template <typename T, int MR>
__device__ __forceinline__ void sbpart_body(const SolveRowTask& t, const T* __restrict__ L,
                                            const T* __restrict__ src, T* __restrict__ part) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T acc[4] = {T(0), T(0), T(0), T(0)};
    if constexpr (MR == 16) {
#pragma unroll 1
        for (int h = 0; h < 2; ++h) sbpart_rows<T, 8>(t, L, src, acc, 8 * h);
    } else {
        sbpart_rows<T, MR>(t, L, src, acc, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int c = wave + 4 * j;
        const T a = wave_sum(acc[j]);
        if (lane == 0 && c < t.nc) part[c] = a;
    }
}

// Backward partial sums: part[d_off + c] = sum over rows [r0, r0 + nr) of L(row, c) x(row) for the
// task's nc (<= 16) columns; one wave per column (lanes over rows), x held in registers.  The
// register tiles cover the task's rows in 2, 4, 8 or 16 chunks of 64 (uniform per workgroup).
template <typename T>
__global__ __launch_bounds__(256) void k_sbpart(const SolveRowTask* __restrict__ tasks, int64_t ntasks,
                                                const T* __restrict__ L,
                                                const T* __restrict__ xall, int64_t ldx,
                                                const T* __restrict__ uall, int64_t ldu,
                                                T* __restrict__ partall, int64_t ldp) {
    const int64_t ti = xcd_chunk(blockIdx.x, gridDim.x);
    if (ti >= ntasks) return;                  // the whole workgroup leaves together
    const SolveRowTask t = tasks[ti];
    if (t.nc < 1) return;
    const T* src = (t.mode == 0 ? xall + blockIdx.y * ldx : uall + blockIdx.y * ldu) + t.v_off;
    T* part = partall + blockIdx.y * ldp + t.d_off;
    const int mr = (t.nr + 63) / 64;
    if (mr <= 2) sbpart_body<T, 2>(t, L, src, part);
    else if (mr <= 4) sbpart_body<T, 4>(t, L, src, part);
    else if (mr <= 8) sbpart_body<T, 8>(t, L, src, part);
    else sbpart_body<T, SOLVE_RC / 64>(t, L, src, part);
}

// Column passes of the backward diagonal (NW waves): a wave's CP columns c + NW j per memory round
// trip (CP = 4, or 2 for 16 row chunks, which keeps the kernel's registers at the 8-chunk size),
// the rows in MZ chunks of 64; rows beyond nz read row 0 (in bounds), weighted by 0.
template <typename T, int MZ, int NW>
__device__ __forceinline__ void sbdiag_cols(const T* __restrict__ X, int nb, int c0, int ncol, int nz,
                                            const T* __restrict__ z, T* __restrict__ x) {
    constexpr int CP = MZ == 16 ? 2 : 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = wave; c < ncol; c += CP * NW) {
        T v[CP][MZ];
#pragma unroll
        for (int j = 0; j < CP; ++j) {
            const T* Xc = X + (int64_t)(c0 + min(c + NW * j, ncol - 1)) * nb + c0;
#pragma unroll
            for (int m = 0; m < MZ; ++m) v[j][m] = Xc[lane + 64 * m < nz ? lane + 64 * m : 0];
        }
#pragma unroll
        for (int j = 0; j < CP; ++j) {
            T acc = T(0);
#pragma unroll
            for (int m = 0; m < MZ; ++m) acc += v[j][m] * ((lane + 64 * m < nz) ? z[lane + 64 * m] : T(0));
            acc = wave_sum(acc);
            if (lane == 0 && c + NW * j < ncol) x[c0 + c + NW * j] = acc;
        }
    }
}

// Backward diagonal: x(columns of tile t) = X(rows >= 64t, columns)^T z, z = y_b - sum of the
// block's partial sums (fixed order); one wave per column, lanes over rows.
template <typename T>
__global__ __launch_bounds__(256) void k_sbdiag(const SolveTile* __restrict__ tiles,
                                                    const SolveBlock* __restrict__ blocks,
                                                    const T* __restrict__ Xb,
                                                    const T* __restrict__ yall, T* __restrict__ xall,
                                                    int64_t ldx, const T* __restrict__ partall, int64_t ldp,
                                                    int sb) {
    const SolveTile st = tiles[blockIdx.x];
    const SolveBlock b = blocks[st.blk];
    const T* y = yall + blockIdx.y * ldx + b.x_off;
    T* x = xall + blockIdx.y * ldx + b.x_off;
    const T* part = partall + blockIdx.y * ldp;
    __shared__ T z[1024];
    constexpr int NW = 4;
    const int tid = threadIdx.x;
    const int c0 = TS * st.t, ncol = min(TS, b.nb - c0), nz = b.nb - c0;
    for (int r = tid; r < nz; r += 64 * NW) {
        // the partial-sum slots are subtracted in slot order, 8 per memory round trip (past the
        // last slot the loads repeat it and are skipped)
        T v = y[c0 + r];
        for (int p = b.part_begin; p < b.part_end; p += 8) {
            T w[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) w[u] = part[(int64_t)min(p + u, b.part_end - 1) * sb + c0 + r];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (p + u < b.part_end) v -= w[u];
        }
        z[r] = v;
    }
    __syncthreads();
    const int mz = (nz + 63) / 64;
    if (mz <= 2) sbdiag_cols<T, 2, NW>(Xb + b.xb_off, b.nb, c0, ncol, nz, z, x);
    else if (mz <= 4) sbdiag_cols<T, 4, NW>(Xb + b.xb_off, b.nb, c0, ncol, nz, z, x);
    else if (mz <= 8) sbdiag_cols<T, 8, NW>(Xb + b.xb_off, b.nb, c0, ncol, nz, z, x);
    else sbdiag_cols<T, 16, NW>(Xb + b.xb_off, b.nb, c0, ncol, nz, z, x);
}

// Backward diagonal, quarter tiles: 16 of the tile's 64 columns per workgroup, one column per
// wave (all its rows in one round trip); z as in k_sbdiag.
template <typename T, int MZ>
__device__ __forceinline__ void sbdiag_col(const T* __restrict__ X, int nb, int c0, int c, int nz,
                                           const T* __restrict__ z, T* __restrict__ x) {
    const int lane = threadIdx.x & 63;
    const T* Xc = X + (int64_t)(c0 + c) * nb + c0;
    T v[MZ];
#pragma unroll
    for (int m = 0; m < MZ; ++m) v[m] = Xc[lane + 64 * m < nz ? lane + 64 * m : 0];
    T acc = T(0);
#pragma unroll
    for (int m = 0; m < MZ; ++m) acc += v[m] * ((lane + 64 * m < nz) ? z[lane + 64 * m] : T(0));
    acc = wave_sum(acc);
    if (lane == 0) x[c0 + c] = acc;
}

template <typename T>
__global__ __launch_bounds__(1024) void k_sbdiag_q(const SolveTile* __restrict__ tiles,
                                                   const SolveBlock* __restrict__ blocks,
                                                   const T* __restrict__ Xb,
                                                   const T* __restrict__ yall, T* __restrict__ xall,
                                                   int64_t ldx, const T* __restrict__ partall, int64_t ldp,
                                                   int sb) {
    const SolveTile st = tiles[blockIdx.x >> 2];
    const SolveBlock b = blocks[st.blk];
    const int c0 = TS * st.t, nz = b.nb - c0, cq = 16 * (blockIdx.x & 3);
    const int ncol = min(16, min(TS, b.nb - c0) - cq);
    if (ncol <= 0) return;                     // the whole workgroup leaves together
    const T* y = yall + blockIdx.y * ldx + b.x_off;
    T* x = xall + blockIdx.y * ldx + b.x_off;
    const T* part = partall + blockIdx.y * ldp;
    __shared__ T z[1024];
    const int tid = threadIdx.x, wave = tid >> 6;
    for (int r = tid; r < nz; r += 1024) {
        T v = y[c0 + r];
        for (int p = b.part_begin; p < b.part_end; p += 8) {
            T w[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) w[u] = part[(int64_t)min(p + u, b.part_end - 1) * sb + c0 + r];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (p + u < b.part_end) v -= w[u];
        }
        z[r] = v;
    }
    __syncthreads();
    if (wave >= ncol) return;
    const int mz = (nz + 63) / 64, c = cq + wave;
    if (mz <= 2) sbdiag_col<T, 2>(Xb + b.xb_off, b.nb, c0, c, nz, z, x);
    else if (mz <= 4) sbdiag_col<T, 4>(Xb + b.xb_off, b.nb, c0, c, nz, z, x);
    else if (mz <= 8) sbdiag_col<T, 8>(Xb + b.xb_off, b.nb, c0, c, nz, z, x);
    else sbdiag_col<T, 16>(Xb + b.xb_off, b.nb, c0, c, nz, z, x);
}

// Grid for an XCD-contiguous task map (xcd_chunk): n workgroups rounded up to whole rounds of the
// 8 XCDs; the kernels skip chunks past their task count.
static unsigned xcd_grid(int64_t n) { return (unsigned)(8 * ((n + 7) / 8)); }

// Block-inverse entries a range of diagonal-solve tiles reads: forward, rows of tile t times the
// columns up to its end; backward, columns of tile t times the rows from its start.
static double diag_tile_entries(const std::vector<SolveTile>& v, const std::vector<SolveBlock>& blocks, int64_t b,
                                int64_t e, bool forward) {
    double n = 0.0;
    for (int64_t q = b; q < e; ++q) {
        const int nb = blocks[v[q].blk].nb, t0 = 64 * v[q].t, w = std::min(64, nb - t0);
        n += (double)w * (forward ? std::min(nb, t0 + 64) : nb - t0);
    }
    return n;
}

// Factor entries a range of solve row tasks reads (nr x nc each).
static double row_task_bytes(const std::vector<SolveRowTask>& v, int64_t b, int64_t e) {
    double n = 0.0;
    for (int64_t q = b; q < e; ++q) n += (double)v[q].nr * v[q].nc;
    return n;
}

// Solve steps of fewer tiles or row tasks than this (per right-hand side) run the quarter-tile
// kernels: with at most a few workgroups per CU the step's time is one workgroup's latency, which
// spreading its bytes over more CUs shortens.  The choice depends on the plan only, so results do
// not depend on the number of right-hand sides.
constexpr int64_t kSolveWideBelow = 512;

// factor-order vector conversions around a scaled (or lower-precision) solve: y = (T)(x * s)
template <typename T>
__global__ void k_to_factor(const double* __restrict__ x, const double* __restrict__ s, T* __restrict__ y,
                            int64_t n, int64_t ld) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = x[blockIdx.y * ld + i];
    y[blockIdx.y * ld + i] = (T)(s ? v * s[i] : v);
}
template <typename T>
__global__ void k_from_factor(const T* __restrict__ y, const double* __restrict__ s, double* __restrict__ x,
                              int64_t n, int64_t ld) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = (double)y[blockIdx.y * ld + i];
    x[blockIdx.y * ld + i] = s ? v * s[i] : v;
}

// ------------------------------------------------------------------------------------------
// Permutation / pin maps, SpMV, norms, refinement helpers
// ------------------------------------------------------------------------------------------
__global__ void k_gather_map(const int64_t* __restrict__ map, int64_t n, const double* __restrict__ src,
                             int64_t lds, double* __restrict__ dst, int64_t ldd) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t t = map[i];
    if (t >= 0) dst[blockIdx.y * ldd + t] = src[blockIdx.y * lds + i];
}
__global__ void k_scatter_map(const int64_t* __restrict__ map, int64_t n, const double* __restrict__ src,
                              int64_t lds, double* __restrict__ dst, int64_t ldd) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t t = map[i];
    dst[blockIdx.y * ldd + i] = (t >= 0) ? src[blockIdx.y * lds + t] : 0.0;
}
template <typename T>
__global__ void k_copy(const T* __restrict__ s, T* __restrict__ d, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

__global__ __launch_bounds__(256) void k_spmv(const int64_t* __restrict__ ptr,
                                              const int64_t* __restrict__ ent,
                                              const int32_t* __restrict__ col,
                                              const double* __restrict__ Ax,
                                              const double* __restrict__ x, double* __restrict__ y,
                                              int64_t nb) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nb) return;
    double y0 = 0, y1 = 0, y2 = 0;
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) {
        const int64_t e = ent[p];
        const int mode = (int)(e & 3);
        const double* B = Ax + 9 * (e >> 2);
        const double* xj = x + 3 * (int64_t)col[p];
        const double x0 = xj[0], x1 = xj[1], x2 = xj[2];
        if (mode == 0) {
            y0 += B[0] * x0 + B[3] * x1 + B[6] * x2;
            y1 += B[1] * x0 + B[4] * x1 + B[7] * x2;
            y2 += B[2] * x0 + B[5] * x1 + B[8] * x2;
        } else if (mode == 1) {
            y0 += B[0] * x0 + B[1] * x1 + B[2] * x2;
            y1 += B[3] * x0 + B[4] * x1 + B[5] * x2;
            y2 += B[6] * x0 + B[7] * x1 + B[8] * x2;
        } else {   // symmetric from the upper triangle: (0,0)=B0 (0,1)=B3 (1,1)=B4 (0,2)=B6 (1,2)=B7 (2,2)=B8
            y0 += B[0] * x0 + B[3] * x1 + B[6] * x2;
            y1 += B[3] * x0 + B[4] * x1 + B[7] * x2;
            y2 += B[6] * x0 + B[7] * x1 + B[8] * x2;
        }
    }
    y[3 * i] = y0;
    y[3 * i + 1] = y1;
    y[3 * i + 2] = y2;
}

__global__ void k_residual(const double* __restrict__ b, const double* __restrict__ y,
                           const int64_t* __restrict__ mask, double* __restrict__ r, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    r[i] = mask[i] >= 0 ? b[i] - y[i] : 0.0;
}
__global__ void k_axpy1(double* __restrict__ x, const double* __restrict__ d, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

constexpr int NRED = 256;
__global__ __launch_bounds__(256) void k_sumsq_partial(const double* __restrict__ a,
                                                       const int64_t* __restrict__ mask, int64_t n,
                                                       double* __restrict__ part) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)NRED * 256)
        if (!mask || mask[i] >= 0) acc += a[i] * a[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}
__global__ void k_sum_final(const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double s[256];
    s[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

// per-segment sum of squares over the free rows (mask >= 0): one workgroup per segment, strided
// partial sums and a fixed-order tree, so the result does not depend on scheduling
__global__ __launch_bounds__(256) void k_seg_sumsq(const double* __restrict__ a, const int64_t* __restrict__ mask,
                                                   const int64_t* __restrict__ seg, double* __restrict__ out) {
    __shared__ double sh[256];
    const int64_t b = seg[blockIdx.x], e = seg[blockIdx.x + 1];
    double acc = 0.0;
    for (int64_t i = b + threadIdx.x; i < e; i += 256)
        if (mask[i] >= 0) acc += a[i] * a[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// trs_kernels::MultiplyLowerBlock semantics on the device: the forward update's row dot
// (lanes over rows, columns in order) with the rows scattered through `indices`.
__global__ void k_mlb(int BS, const int64_t* __restrict__ indices, int64_t start, int64_t size,
                      int64_t degree, const double* __restrict__ A, int64_t lda,
                      double* __restrict__ B, int64_t ldb) {
    extern __shared__ double y[];
    double* Bc = B + blockIdx.y * ldb;
    for (int64_t c = threadIdx.x; c < size; c += blockDim.x) y[c] = Bc[start + c];
    __syncthreads();
    for (int64_t i = threadIdx.x; i < degree; i += blockDim.x) {
        double v = 0.0;
        for (int64_t c = 0; c < size; ++c) v += A[c * lda + i] * y[c];
        Bc[indices[i / BS] + i % BS] -= v;
    }
}

// ------------------------------------------------------------------------------------------
// FactorizerDev
// ------------------------------------------------------------------------------------------
FactorizerDev::FactorizerDev(int prec) : precision(prec) {
    if (prec < 0 || prec > 3) throw Error(-1, "unknown precision");
    // persistent grid of the fused level kernel: two workgroups per compute unit
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    fused_grid_ = 2 * std::max(ncu, 1);
}
FactorizerDev::~FactorizerDev() {
    drop_graph();
    for (void* e : join_)
        if (e) (void)hipEventDestroy((hipEvent_t)e);
    for (void* e : la_ev_) (void)hipEventDestroy((hipEvent_t)e);
    if (cstream_) (void)hipStreamDestroy((hipStream_t)cstream_);
    if (side_) (void)hipStreamDestroy((hipStream_t)side_);
}

void FactorizerDev::require(int t) const {
    if (ftype < t)
        throw Error(-2, t >= 2 ? "Numeric factorization does not exist" : "Symbolic factorization does not exist");
}

void FactorizerDev::ensure_vec(DeviceArray& a, size_t bytes) { a.alloc(bytes); }

size_t FactorizerDev::fsize() const { return precision != 0 ? sizeof(float) : sizeof(double); }

void FactorizerDev::symbolic(const int64_t* Ap, const int64_t* Ai, int64_t nb, int bs,
                             const int64_t* pins, int64_t npins) {
    if (bs != 3) throw Error(-6, "only uniform 3x3 blocks are supported");
    SymbolicOptions o = options_from_env();
    o.nd_state = ordering_state(pins, npins);
    o.schur_tile = schur_tile();
    install(std::make_unique<Symbolic>(analyze(Ap, Ai, nb, pins, npins, o)));
}

void FactorizerDev::symbolic_scalar(const int64_t* Ap, const int64_t* Ai, int64_t n, const int64_t* pins,
                                    int64_t npins) {
    ScalarBlocking B = block_scalar_pattern(Ap, Ai, n);
    SymbolicOptions o = options_from_env();
    o.nd_state = ordering_state(pins, npins);
    o.schur_tile = schur_tile();
    o.n_scalar = n;
    install(std::make_unique<Symbolic>(analyze(B.P.Ap.data(), B.P.Ai.data(), B.P.n, pins, npins, o)));
    scalar_mode_ = true;
    scalar_nnz_ = (int64_t)B.value_pos.size();
    s2b_.upload(B.value_pos);
    Axb_.alloc(std::max<int64_t>(9 * B.P.nz(), 1) * sizeof(double));
    Bxb_.release();
}

__global__ void k_scalar_to_block(const int64_t* __restrict__ pos, int64_t n, const double* __restrict__ a,
                                  double* __restrict__ ab) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n) ab[pos[k]] = a[k];
}

void FactorizerDev::scalar_values(const double* Ax, double* Axb, void* stream) {
    hipStream_t s = S_(stream);
    fill<double>(Axb, 9 * sym->nz_full, 0.0, s);
    if (scalar_nnz_) {
        k_scalar_to_block<<<nblk(scalar_nnz_, 256), 256, 0, s>>>(s2b_.as<int64_t>(), scalar_nnz_, Ax, Axb);
        LAUNCH_CHECK();
    }
}

void FactorizerDev::numeric_scalar(const double* Ax, double sigma, const double* Bx, void* stream) {
    require(1);
    if (!scalar_mode_) throw Error(-2, "the symbolic factorization is of a block matrix");
    scalar_values(Ax, Axb_.as<double>(), stream);
    const double* bxb = nullptr;
    if (Bx) {
        Bxb_.alloc(Axb_.bytes);
        scalar_values(Bx, Bxb_.as<double>(), stream);
        bxb = Bxb_.as<double>();
    }
    numeric(Axb_.as<double>(), sigma, bxb, stream);
}

// 256x256 Schur tiles for the bfloat16 Schur-update kernel (MFS_BF16_TILE=128 keeps 128x128)
int FactorizerDev::schur_tile() const {
    if (precision < 2) return 128;
    const char* e = getenv("MFS_BF16_TILE");
    return (e && atoi(e) == 128) ? 128 : 256;
}

// The ordering history is kept across analyses with the same pinned variables
// (CatamariFactorizer.hh:236-244: a change of the pins restarts it).
NDState* FactorizerDev::ordering_state(const int64_t* pins, int64_t npins) {
    std::vector<int64_t> sp(pins, pins + npins);
    std::sort(sp.begin(), sp.end());
    if (sp != nd_pins_) nd_state_.reset();
    nd_pins_ = std::move(sp);
    return &nd_state_;
}

void FactorizerDev::set_reuse_period(int64_t period) {
    if (period < 0) throw Error(-1, "the temporal reuse period must be >= 0");
    if (period != nd_state_.period) nd_state_.reset();
    nd_state_.period = period;
}

// Upload a symbolic analysis (task lists, maps) and size the factor storage for it.
void FactorizerDev::install(std::unique_ptr<Symbolic> S) {
    drop_graph();
    ftype = 0;
    scalar_mode_ = false;
    stash_valid_ = false;
    // injection plan entries with per-entry leading dimension and factor-order block row/column
    std::vector<PlanEnt> pe(S->plan_src.size());
    for (size_t q = 0; q < pe.size(); ++q) {
        const int64_t dst = S->plan_dst[q] >> 2;
        const int32_t s = (int32_t)(std::upper_bound(S->panel_off.begin(), S->panel_off.end(), dst) -
                                    S->panel_off.begin()) - 1;
        pe[q].dstmode = S->plan_dst[q];
        pe[q].src = S->plan_src[q];
        pe[q].ld = 3 * (S->swidth[s] + S->sdeg[s]);
        pe[q].rb = S->plan_row[q];
        pe[q].cb = S->plan_col[q];
    }
    planent_.upload(pe);
    diagsrc_.upload(S->diag_src);
    rel_.upload(S->rel);
    ext_.upload(S->extend_tasks);
    potrf_.upload(S->potrf_tasks);
    panel_.upload(S->panel_items);
    // shared-tile counters start at zero and every fused launch leaves them at zero
    pcount_.alloc(std::max<size_t>(S->potrf_tasks.size(), 1) * sizeof(int));
    HIP_CHECK(hipMemset(pcount_.p, 0, pcount_.bytes));
    fused_.upload(S->fused_tasks);
    uoff_.upload(S->u_off);
    gemm_.upload(S->gemm_tasks);
    pull_.upload(S->pull_ents);
    ppull_.upload(S->panel_pull_tasks);
    tinj_.upload(S->tile_inj);
    fillt_.upload(S->fill_tiles);
    gath_.upload(S->gather_tasks);
    sblk_.upload(S->sblocks);
    invt_.upload(S->inv_tiles);
    fdt_.upload(S->fdiag_tiles);
    bdt_.upload(S->bdiag_tiles);
    fupd_.upload(S->fupd_tasks);
    bprt_.upload(S->bpart_tasks);
    tslot_.upload(S->tile_slot);
    children_.upload(S->children);
    usoff_.upload(S->u_solve_off);
    rows_.upload(S->rows);
    sdeg_.upload(S->sdeg);
    reloff_.upload(S->rel_off);
    spmv_ptr_.upload(S->spmv_ptr);
    spmv_ent_.upload(S->spmv_ent);
    spmv_col_.upload(S->spmv_col);
    f2p_.upload(S->full_to_perm);
    f2r_.upload(S->full_to_red);
    f2up_.upload(S->full_to_uperm);
    r2p_.upload(S->red_to_int);
    up2i_.upload(S->uperm_to_int);
    if (S->masked()) cmask_.upload(S->comp_mask);
    else cmask_.release();
    const size_t fs = fsize();
    // (tails: the direct-to-LDS tile GEMM reads up to 128 elements past an operand column's start)
    L_.alloc((std::max<int64_t>(S->L_size, 1) + 256) * fs);
    U_.alloc(std::max<int64_t>(S->U_size, 1) * fs);
    inv_.alloc((S->inv_slots * TS * TS + 256) * fs);
    Xb_.alloc(S->xb_size * fs);
    if (precision != 0) scale_.alloc(3 * S->n_red * sizeof(double));
    status_.alloc(sizeof(int) * 4);
    sym = std::move(S);
    ftype = 1;
}

// Every device operation of one numeric factorization, enqueued on `s` (also the body of the
// captured graph).  ev_inject, if given, is recorded after the value injection.
template <typename T>
void FactorizerDev::enqueue_numeric(const double* Ax, double sigma, const double* Bx, void* stream,
                                    void* ev_inject) {
    const Symbolic& S = *sym;
    hipStream_t s = S_(stream);
    T* L = L_.as<T>();
    T* U = U_.as<T>();
    T* inv = inv_.as<T>();
    prof.begin(KernelProfile::FILL_L, s);
    if (panel_pulls_active()) {
        // the panel tiles with children lists are rebuilt whole at their level: clear the rest
        if (!S.fill_tiles.empty()) {
            k_fill_tiles<T><<<(unsigned)S.fill_tiles.size(), 256, 0, s>>>(fillt_.as<FillTile>(), L);
            LAUNCH_CHECK();
        }
    } else {
        fill<T>(L, S.L_size, T(0), s);
    }
    prof.end(s);
    k_set_int<<<1, 1, 0, s>>>(status_.as<int>(), INT_MAX);
    LAUNCH_CHECK();
    const double* scale = nullptr;
    if (precision != 0) {   // symmetric diagonal scaling of the factored matrix
        k_diag_scale<<<nblk(S.n_red, 256), 256, 0, s>>>(diagsrc_.as<int32_t>(), S.n_red, Ax, Bx, sigma,
                                                       cmask(), scale_.as<double>(), status_.as<int>());
        LAUNCH_CHECK();
        scale = scale_.as<double>();
    }
    const int64_t np = (int64_t)S.plan_src.size();
    if (np) {
        prof.begin(KernelProfile::INJECT, s);
        k_inject<T><<<nblk(np, 256), 256, 0, s>>>(planent_.as<PlanEnt>(), np, Ax, Bx, sigma, scale, cmask(), L);
        LAUNCH_CHECK();
        prof.end(s);
    }
    if (ev_inject) HIP_CHECK(hipEventRecord((hipEvent_t)ev_inject, s));
    const GemmTask* gt = gemm_.as<GemmTask>();
    // the Schur-update tiles add the children's update-matrix blocks in their epilogue (in place
    // of the extend-add passes into U) where every tile carries its list
    const bool pulls = pulls_active();
    // Schur updates (in-panel and update matrix) of the BF16 variants go through the bfloat16
    // tile kernel; TRSM and the diagonal tiles stay in the factor type
    // (`syrk`: the range is a Schur update U = -L21 L21^T, whose large tiles are S.schur_tile)
    auto schur_gemm = [&](const TaskRange& r, hipStream_t st, bool syrk = false) {
        if constexpr (sizeof(T) == 4) {
            if (precision >= 2) {
                launch_gemm_bf(r, gt, (float*)L, (float*)U, (const float*)inv, precision == 3,
                               syrk && S.schur_tile == 256, st, syrk && pulls ? pull_.as<PullEnt>() : nullptr,
                               rel_.as<int32_t>());
                return;
            }
        }
        if (syrk && pulls) launch_gemm<T>(r, gt, L, U, inv, st, pull_.as<PullEnt>(), rel_.as<int32_t>());
        else launch_gemm<T>(r, gt, L, U, inv, st);
    };
    // lookahead: the rest of an outer trailing update runs on the side stream after the step's
    // TRSM; the stream waits for it before the next outer trailing update or the Schur update,
    // which write the same columns (so every entry still receives its updates in block order)
    hipStream_t side = nullptr;
    size_t nev = 0;
    hipEvent_t pending = nullptr;
    auto event = [&]() {
        if (nev == la_ev_.size()) {
            hipEvent_t e;
            HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            la_ev_.push_back(e);
        }
        return (hipEvent_t)la_ev_[nev++];
    };
    auto join = [&]() {
        if (pending) HIP_CHECK(hipStreamWaitEvent(s, pending, 0));
        pending = nullptr;
    };
    auto extend = [&](const std::vector<std::pair<int64_t,int64_t>>& passes) {
        for (auto [a, b] : passes)
            if (b > a) {
                prof.begin(KernelProfile::EXTEND, s);
                k_extend<T><<<(unsigned)((b - a + 3) / 4), 256, 0, s>>>(ext_.as<ExtendTask>() + a, b - a, rel_.as<int32_t>(), U, L);
                LAUNCH_CHECK();
                prof.end(s);
            }
    };
    for (const LevelPlan& P : S.levels) {
        const int64_t nf = P.fused.second - P.fused.first;
        if (nf > 0) {
            prof.begin(KernelProfile::FUSED, s, 0.0, nf);
            k_fused<T><<<(unsigned)std::min<int64_t>(nf, fused_grid_), 256, 0, s>>>(
                fused_.as<FusedTask>() + P.fused.first, (int)nf, children_.as<int32_t>(), sdeg_.as<int32_t>(),
                reloff_.as<int64_t>(), rel_.as<int32_t>(), uoff_.as<int64_t>(), L, U, inv, status_.as<int>());
            LAUNCH_CHECK();
            prof.end(s);
        }
        if (panel_pulls_active()) {
            const int64_t n = P.panel_pulls.second - P.panel_pulls.first;
            if (n > 0) {
                prof.begin(KernelProfile::EXTEND, s, 0.0, n);
                k_panel_pull<T><<<(unsigned)n, 256, 0, s>>>(
                    ppull_.as<PanelPullTask>() + P.panel_pulls.first, pull_.as<PullEnt>(), rel_.as<int32_t>(),
                    tinj_.as<TileInj>(), planent_.as<PlanEnt>(), Ax, Bx, sigma, scale, cmask(), L, U);
                LAUNCH_CHECK();
                prof.end(s);
            }
        } else {
            extend(P.extend_passes);
        }
        for (size_t t = 0; t < P.potrf_steps.size(); ++t) {
            auto [p0, p1] = P.potrf_steps[t];
            const TaskRange& tr = P.trsm_steps[t];
            auto [i0, i1] = P.panel_steps[t];
            if (i1 > i0) {
                prof.begin(KernelProfile::PANEL, s, prof.on ? gemm_work(S, tr) : 0.0, i1 - i0);
                k_panel<T><<<(unsigned)(i1 - i0), 256, 0, s>>>(panel_.as<PanelItem>() + i0, potrf_.as<PotrfTask>(),
                                                             gt, L, inv, status_.as<int>(), pcount_.as<int>());
                LAUNCH_CHECK();
                prof.end(s);
            } else if (p1 > p0) {
                prof.begin(KernelProfile::POTRF, s, 0.0, p1 - p0);
                k_potrf<T><<<(unsigned)(p1 - p0), 256, 0, s>>>(potrf_.as<PotrfTask>() + p0, L, inv, status_.as<int>());
                LAUNCH_CHECK();
                prof.end(s);
            }
            if (tr.e > tr.b && i1 == i0) {
                prof.begin(KernelProfile::TRSM, s, prof.on ? gemm_work(S, tr) : 0.0, tr.e - tr.b,
                           prof.on ? gemm_noreuse_bytes(S, tr, sizeof(T)) : 0.0);
                launch_gemm<T>(tr, gt, L, U, inv, s);
                prof.end(s);
            }
            const TaskRange& ur = P.update_steps[t];
            const TaskRange& rr = P.rest_steps[t];
            if (P.trailing_step[t]) join();
            if (rr.e > rr.b && prof.on) {
                // per-kernel timing: the rest runs in order on the stream, so the measured launch
                // durations are not stretched by kernels sharing the device
                prof.begin(KernelProfile::UPDATE, s, gemm_work(S, rr), rr.e - rr.b,
                           gemm_noreuse_bytes(S, rr, sizeof(T)));
                schur_gemm(rr, s);
                prof.end(s);
            } else if (rr.e > rr.b) {
                if (!side_) {
                    hipStream_t ss;
                    HIP_CHECK(hipStreamCreateWithFlags(&ss, hipStreamNonBlocking));
                    side_ = ss;
                }
                side = (hipStream_t)side_;
                hipEvent_t e = event();
                HIP_CHECK(hipEventRecord(e, s));
                HIP_CHECK(hipStreamWaitEvent(side, e, 0));
                prof.begin(KernelProfile::UPDATE, side, prof.on ? gemm_work(S, rr) : 0.0, rr.e - rr.b,
                           prof.on ? gemm_noreuse_bytes(S, rr, sizeof(T)) : 0.0);
                schur_gemm(rr, side);
                prof.end(side);
                pending = event();
                HIP_CHECK(hipEventRecord(pending, side));
            }
            if (ur.e > ur.b) {
                prof.begin(KernelProfile::UPDATE, s, prof.on ? gemm_work(S, ur) : 0.0, ur.e - ur.b,
                           prof.on ? gemm_noreuse_bytes(S, ur, sizeof(T)) : 0.0);
                schur_gemm(ur, s);
                prof.end(s);
            }
        }
        join();
        if (P.syrk.e > P.syrk.b) {
            prof.begin(KernelProfile::SYRK, s, prof.on ? gemm_work(S, P.syrk) : 0.0, P.syrk.e - P.syrk.b,
                       prof.on ? gemm_noreuse_bytes(S, P.syrk, sizeof(T)) : 0.0);
            schur_gemm(P.syrk, s, true);
            prof.end(s);
        }
        if (!pulls) extend(P.extend_u_passes);
    }
    // explicit inverses of the solve column blocks' diagonal parts
    if (!S.inv_tiles.empty()) {
        prof.begin(KernelProfile::BLKINV, s, 0.0, (int64_t)S.inv_tiles.size());
        k_blkinv<T><<<(unsigned)S.inv_tiles.size(), 256, 0, s>>>(invt_.as<SolveTile>(), sblk_.as<SolveBlock>(),
                                                               tslot_.as<int32_t>(), L, inv, Xb_.as<T>());
        LAUNCH_CHECK();
        prof.end(s);
    }
}

void FactorizerDev::drop_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy((hipGraphExec_t)graph_exec_);
    if (graph_) (void)hipGraphDestroy((hipGraph_t)graph_);
    graph_exec_ = graph_ = nullptr;
    graph_nodes_ = 0;
}

void FactorizerDev::set_graph(bool on) {
    graph_on_ = on;
    if (!on) drop_graph();
}

template <typename T>
void FactorizerDev::numeric_t(const double* Ax, double sigma, const double* Bx, void* stream) {
    const Symbolic& S = *sym;
    hipStream_t s = S_(stream);
    hipEvent_t ev[3];
    for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipEventRecord(ev[0], s));
    if (graph_on_) {
        // the whole factorization as one captured graph on the library's own stream (a graph
        // cannot be captured on the legacy default stream), joined to the caller's stream by
        // events; re-captured when the arguments or the schedule change
        if (!cstream_) {
            hipStream_t cs;
            HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            cstream_ = cs;
            for (auto& e : join_) {
                hipEvent_t je;
                HIP_CHECK(hipEventCreateWithFlags(&je, hipEventDisableTiming));
                e = je;
            }
        }
        hipStream_t cs = (hipStream_t)cstream_;
        HIP_CHECK(hipEventRecord((hipEvent_t)join_[0], s));
        HIP_CHECK(hipStreamWaitEvent(cs, (hipEvent_t)join_[0], 0));
        if (!graph_exec_ || gkey_ax_ != Ax || gkey_bx_ != Bx || gkey_sigma_ != sigma) {
            drop_graph();
            const bool prof_on = prof.on;
            prof.on = false;            // no event records inside the capture
            HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            hipGraph_t g = nullptr;
            try {
                enqueue_numeric<T>(Ax, sigma, Bx, cs, nullptr);
            } catch (...) {
                (void)hipStreamEndCapture(cs, &g);
                if (g) (void)hipGraphDestroy(g);
                prof.on = prof_on;
                throw;
            }
            prof.on = prof_on;
            HIP_CHECK(hipStreamEndCapture(cs, &g));
            graph_ = g;
            hipGraphExec_t ge;
            HIP_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            graph_exec_ = ge;
            size_t nn = 0;
            HIP_CHECK(hipGraphGetNodes(g, nullptr, &nn));
            graph_nodes_ = (int64_t)nn;
            gkey_ax_ = Ax;
            gkey_bx_ = Bx;
            gkey_sigma_ = sigma;
        }
        HIP_CHECK(hipGraphLaunch((hipGraphExec_t)graph_exec_, cs));
        HIP_CHECK(hipEventRecord((hipEvent_t)join_[1], cs));
        HIP_CHECK(hipStreamWaitEvent(s, (hipEvent_t)join_[1], 0));
        HIP_CHECK(hipEventRecord(ev[1], s));
    } else {
        enqueue_numeric<T>(Ax, sigma, Bx, stream, ev[1]);
    }
    HIP_CHECK(hipEventRecord(ev[2], s));
    int st = 0;
    HIP_CHECK(hipMemcpyAsync(&st, status_.as<int>(), sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    prof.collect();
    float m01 = 0, m12 = 0;
    HIP_CHECK(hipEventElapsedTime(&m01, ev[0], ev[1]));
    HIP_CHECK(hipEventElapsedTime(&m12, ev[1], ev[2]));
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (graph_on_) std::swap(m01, m12);   // one graph launch: no injection split
    timings[0] = m01;
    timings[1] = 0;
    timings[2] = m12;
    timings[3] = m01 + m12;
    if (st != INT_MAX) {
        // pivots before the failing one, counted over the free scalars (masked ones are the identity)
        num_successful = S.masked() ? (int64_t)(std::lower_bound(S.uperm_to_int.begin(), S.uperm_to_int.end(),
                                                                 (int64_t)st) - S.uperm_to_int.begin())
                                    : st;
        ftype = 1;
        throw Error(-3, std::to_string(num_successful) + "/" + std::to_string(S.n_free) +
                            "  pivots successful in numeric factorization (non-positive definite?)");
    }
    num_successful = S.n_free;
    ftype = 2;
}

void FactorizerDev::numeric(const double* Ax, double sigma, const double* Bx, void* stream) {
    require(1);
    ftype = 1;
    if (precision != 0) numeric_t<float>(Ax, sigma, Bx, stream);
    else numeric_t<double>(Ax, sigma, Bx, stream);
}

// Supernodal solves in factor order, in place on xp (nrhs columns, leading dimension 3 n_red).
// Forward, level by level: gather the children's update vectors, then per column-block step
// y_b = X_b x_b (into Y) and the update of every row below.  Backward, top level first: gather
// the final x of each supernode's structure rows, then per step (last block first) partial sums
// of L(rows below, b)^T x and x_b = X_b^T (y_b - partials) (into xp).
template <typename T>
void FactorizerDev::solve_t(T* xp, int nrhs, void* stream, const SolveScratch* ws) {
    const Symbolic& S = *sym;
    hipStream_t s = S_(stream);
    const int64_t ldx = 3 * S.n_red;
    const int64_t ldu = std::max<int64_t>(S.u_solve_size, 1);
    const int64_t ldp = S.bpart_size * S.solve_block;
    const size_t fs = sizeof(T);
    T *u, *part, *Y;
    if (ws) {
        u = static_cast<T*>(ws->u);
        part = static_cast<T*>(ws->part);
        Y = static_cast<T*>(ws->Y);
    } else {
        usolve_.alloc((size_t)ldu * nrhs * fs);
        bpart_.alloc((size_t)ldp * nrhs * fs);
        ysolve_.alloc((size_t)ldx * nrhs * fs);
        u = usolve_.as<T>();
        part = bpart_.as<T>();
        Y = ysolve_.as<T>();
    }
    const T* L = L_.as<T>();
    const T* Xb = Xb_.as<T>();
    const SolveBlock* blk = sblk_.as<SolveBlock>();
    const SolveGatherTask* gt = gath_.as<SolveGatherTask>();
    for (const LevelPlan& P : S.levels) {
        if (P.gather.second > P.gather.first) {
            dim3 g((unsigned)(P.gather.second - P.gather.first), (unsigned)nrhs);
            prof.begin(KernelProfile::FGATHER, s);
            k_fgather<T><<<g, 256, 0, s>>>(gt + P.gather.first, children_.as<int32_t>(), sdeg_.as<int32_t>(),
                                           reloff_.as<int64_t>(), rel_.as<int32_t>(), usoff_.as<int64_t>(),
                                           xp, ldx, u, ldu);
            LAUNCH_CHECK();
            prof.end(s);
        }
        for (size_t t = 0; t < P.fdiag_steps.size(); ++t) {
            auto [d0, d1] = P.fdiag_steps[t];
            if (d1 > d0) {
                dim3 g((unsigned)(d1 - d0), (unsigned)nrhs);
                prof.begin(KernelProfile::FDIAG, s, 0.0, 0,
                           prof.on ? diag_tile_entries(S.fdiag_tiles, S.sblocks, d0, d1, true) * sizeof(T) * nrhs : 0.0);
                if (d1 - d0 < kSolveWideBelow)
                    k_sfdiag_q<T><<<dim3(4 * g.x, g.y), 1024, 0, s>>>(fdt_.as<SolveTile>() + d0, blk, Xb, xp, Y, ldx);
                else
                    k_sfdiag<T><<<g, 256, 0, s>>>(fdt_.as<SolveTile>() + d0, blk, Xb, xp, Y, ldx);
                LAUNCH_CHECK();
                prof.end(s);
            }
            auto [u0, u1] = P.fupd_steps[t];
            if (u1 > u0) {
                const int64_t nt = u1 - u0;
                prof.begin(KernelProfile::FGEMV, s, 0.0, 0, prof.on ? row_task_bytes(S.fupd_tasks, u0, u1) * sizeof(T) * nrhs : 0.0);
                if (nt < kSolveWideBelow)
                    k_sfupd_q<T><<<dim3(xcd_grid(4 * nt), (unsigned)nrhs), 1024, 0, s>>>(
                        fupd_.as<SolveRowTask>() + u0, nt, L, Y, xp, ldx, u, ldu);
                else
                    k_sfupd<T><<<dim3(xcd_grid(nt), (unsigned)nrhs), 256, 0, s>>>(fupd_.as<SolveRowTask>() + u0, nt,
                                                                                     L, Y, xp, ldx, u, ldu);
                LAUNCH_CHECK();
                prof.end(s);
            }
        }
    }
    for (int lv = (int)S.levels.size() - 1; lv >= 0; --lv) {
        const LevelPlan& P = S.levels[lv];
        if (P.gather.second > P.gather.first) {
            dim3 g((unsigned)(P.gather.second - P.gather.first), (unsigned)nrhs);
            prof.begin(KernelProfile::BGATHER, s);
            k_bgather<T><<<g, 256, 0, s>>>(gt + P.gather.first, rows_.as<int32_t>(), xp, ldx, u, ldu);
            LAUNCH_CHECK();
            prof.end(s);
        }
        for (size_t t = 0; t < P.bdiag_steps.size(); ++t) {
            auto [p0, p1] = P.bpart_steps[t];
            if (p1 > p0) {
                prof.begin(KernelProfile::BGEMV, s, 0.0, 0, prof.on ? row_task_bytes(S.bpart_tasks, p0, p1) * sizeof(T) * nrhs : 0.0);
                k_sbpart<T><<<dim3(xcd_grid(p1 - p0), (unsigned)nrhs), 256, 0, s>>>(bprt_.as<SolveRowTask>() + p0, p1 - p0,
                                                                                    L, xp, ldx, u, ldu, part, ldp);
                LAUNCH_CHECK();
                prof.end(s);
            }
            auto [d0, d1] = P.bdiag_steps[t];
            if (d1 > d0) {
                dim3 g((unsigned)(d1 - d0), (unsigned)nrhs);
                prof.begin(KernelProfile::BDIAG, s, 0.0, 0,
                           prof.on ? diag_tile_entries(S.bdiag_tiles, S.sblocks, d0, d1, false) * sizeof(T) * nrhs : 0.0);
                if (d1 - d0 < kSolveWideBelow)
                    k_sbdiag_q<T><<<dim3(4 * g.x, g.y), 1024, 0, s>>>(bdt_.as<SolveTile>() + d0, blk, Xb, Y, xp, ldx,
                                                                     part, ldp, S.solve_block);
                else
                    k_sbdiag<T><<<g, 256, 0, s>>>(bdt_.as<SolveTile>() + d0, blk, Xb, Y, xp, ldx, part, ldp,
                                                     S.solve_block);
                LAUNCH_CHECK();
                prof.end(s);
            }
        }
    }
}

// factor-order double vectors in/out; FP32 factors solve the scaled system in single precision
void FactorizerDev::solve_permuted_inplace(double* xp, int nrhs, void* stream, const SolveScratch* ws) {
    const Symbolic& S = *sym;
    if (precision == 0) {
        solve_t<double>(xp, nrhs, stream, ws);
        return;
    }
    hipStream_t s = S_(stream);
    const int64_t n = 3 * S.n_red;
    float* xf;
    if (ws) {
        xf = ws->xf;
    } else {
        xf_.alloc((size_t)n * nrhs * sizeof(float));
        xf = xf_.as<float>();
    }
    dim3 g(nblk(n, 256), (unsigned)nrhs);
    prof.begin(KernelProfile::CONVERT, s);
    k_to_factor<float><<<g, 256, 0, s>>>(xp, scale_.as<double>(), xf, n, n);
    LAUNCH_CHECK();
    prof.end(s);
    solve_t<float>(xf, nrhs, stream, ws);
    prof.begin(KernelProfile::CONVERT, s);
    k_from_factor<float><<<g, 256, 0, s>>>(xf, scale_.as<double>(), xp, n, n);
    LAUNCH_CHECK();
    prof.end(s);
}

void FactorizerDev::solve(const double* b, double* x, int nrhs, void* stream) {
    require(2);
    if (nrhs < 1) throw Error(-1, "Must specify at least one rhs.");
    const Symbolic& S = *sym;
    const int64_t ni = 3 * S.n_red;
    scratch_b_.alloc((size_t)ni * nrhs * sizeof(double));
    solve_full_with(b, x, nrhs, scratch_b_.as<double>(), stream);
}

// Bytes of one solve's scratch for nrhs columns, each piece 256-byte aligned: xp, Y, u, part
// (and xf for single-precision factors).
size_t FactorizerDev::solve_scratch_bytes(int nrhs) const {
    const Symbolic& S = *sym;
    const size_t fs = fsize();
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t ni = 3 * (size_t)S.n_red, r = (size_t)nrhs;
    size_t b = up(ni * r * sizeof(double)) + up(ni * r * fs) +
               up((size_t)std::max<int64_t>(S.u_solve_size, 1) * r * fs) +
               up((size_t)std::max<int64_t>(S.bpart_size * S.solve_block, 1) * r * fs);
    if (precision != 0) b += up(ni * r * sizeof(float));
    return b;
}

void FactorizerDev::solve_in(const double* b, double* x, int nrhs, void* base, size_t bytes, void* stream) {
    require(2);
    if (nrhs < 1) throw Error(-1, "Must specify at least one rhs.");
    const size_t need = solve_scratch_bytes(nrhs);
    if (need > bytes)
        throw Error(-1, "the solve needs " + std::to_string(need) + " bytes of scratch for " + std::to_string(nrhs) +
                            " right-hand side(s); the allocation-free solve has " + std::to_string(bytes));
    const Symbolic& S = *sym;
    const size_t fs = fsize();
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    const size_t ni = 3 * (size_t)S.n_red, r = (size_t)nrhs;
    char* p = static_cast<char*>(base);
    SolveScratch ws;
    ws.xp = reinterpret_cast<double*>(p);
    p += up(ni * r * sizeof(double));
    ws.Y = p;
    p += up(ni * r * fs);
    ws.u = p;
    p += up((size_t)std::max<int64_t>(S.u_solve_size, 1) * r * fs);
    ws.part = p;
    p += up((size_t)std::max<int64_t>(S.bpart_size * S.solve_block, 1) * r * fs);
    if (precision != 0) ws.xf = reinterpret_cast<float*>(p);
    solve_full_with(b, x, nrhs, ws.xp, stream, &ws);
}

// the full-length solve through a caller-provided factor-order vector xp (3 n_red x nrhs)
void FactorizerDev::solve_full_with(const double* b, double* x, int nrhs, double* xp, void* stream,
                                    const SolveScratch* ws) {
    const Symbolic& S = *sym;
    hipStream_t s = S_(stream);
    const int64_t nf = S.n_full, ni = 3 * S.n_red;
    if (S.masked()) fill<double>(xp, ni * nrhs, 0.0, s);     // masked positions take no value
    dim3 g(nblk(nf, 256), (unsigned)nrhs);
    k_gather_map<<<g, 256, 0, s>>>(f2p_.as<int64_t>(), nf, b, nf, xp, ni);
    LAUNCH_CHECK();
    solve_permuted_inplace(xp, nrhs, stream, ws);
    k_scatter_map<<<g, 256, 0, s>>>(f2p_.as<int64_t>(), nf, xp, ni, x, nf);
    LAUNCH_CHECK();
}

void FactorizerDev::solve_raw_reduced(const double* b, double* x, int nrhs, bool already_permuted,
                                      void* stream) {
    require(2);
    if (nrhs < 1) throw Error(-1, "Must specify at least one rhs.");
    const Symbolic& S = *sym;
    hipStream_t s = S_(stream);
    const int64_t nr = S.n_free, ni = 3 * S.n_red;
    scratch_b_.alloc((size_t)ni * nrhs * sizeof(double));
    double* xp = scratch_b_.as<double>();
    dim3 g(nblk(nr, 256), (unsigned)nrhs);
    if (already_permuted && !S.masked()) {
        k_copy<double><<<std::min<unsigned>(nblk(nr * nrhs, 256), 8192), 256, 0, s>>>(b, xp, nr * nrhs);
        LAUNCH_CHECK();
    } else {
        if (S.masked()) fill<double>(xp, ni * nrhs, 0.0, s);
        k_gather_map<<<g, 256, 0, s>>>((already_permuted ? up2i_ : r2p_).as<int64_t>(), nr, b, nr, xp, ni);
        LAUNCH_CHECK();
    }
    solve_permuted_inplace(xp, nrhs, stream);
    if (already_permuted && !S.masked()) {
        k_copy<double><<<std::min<unsigned>(nblk(nr * nrhs, 256), 8192), 256, 0, s>>>(xp, x, nr * nrhs);
        LAUNCH_CHECK();
    } else {
        k_scatter_map<<<g, 256, 0, s>>>((already_permuted ? up2i_ : r2p_).as<int64_t>(), nr, xp, ni, x, nr);
        LAUNCH_CHECK();
    }
}

void FactorizerDev::remove_fixed(const double* xf, double* xr, int nrhs, bool permute, void* stream) {
    require(1);
    const Symbolic& S = *sym;
    const int64_t nf = S.n_full, nr = S.n_free;
    dim3 g(nblk(nf, 256), (unsigned)nrhs);
    k_gather_map<<<g, 256, 0, S_(stream)>>>((permute ? f2up_ : f2r_).as<int64_t>(), nf, xf, nf, xr, nr);
    LAUNCH_CHECK();
}

void FactorizerDev::extract_full(const double* xr, double* xf, int nrhs, bool permute, void* stream) {
    require(1);
    const Symbolic& S = *sym;
    const int64_t nf = S.n_full, nr = S.n_free;
    dim3 g(nblk(nf, 256), (unsigned)nrhs);
    k_scatter_map<<<g, 256, 0, S_(stream)>>>((permute ? f2up_ : f2r_).as<int64_t>(), nf, xr, nr, xf, nf);
    LAUNCH_CHECK();
}

void FactorizerDev::spmv(const double* Ax, const double* x, double* y, void* stream) {
    require(1);
    const int64_t nb = sym->n_full_blocks;
    hipStream_t s = S_(stream);
    if (scalar_mode_) {
        // scalar values and length-n vectors: through the block values and padded vectors
        const int64_t n = sym->n_full;
        scalar_values(Ax, Axb_.as<double>(), stream);
        xpad_.alloc(3 * nb * sizeof(double));
        ypad_.alloc(3 * nb * sizeof(double));
        fill<double>(xpad_.as<double>(), 3 * nb, 0.0, s);
        k_copy<double><<<std::min<unsigned>(nblk(n, 256), 8192), 256, 0, s>>>(x, xpad_.as<double>(), n);
        LAUNCH_CHECK();
        k_spmv<<<nblk(nb, 256), 256, 0, s>>>(spmv_ptr_.as<int64_t>(), spmv_ent_.as<int64_t>(),
                                            spmv_col_.as<int32_t>(), Axb_.as<double>(), xpad_.as<double>(),
                                            ypad_.as<double>(), nb);
        LAUNCH_CHECK();
        k_copy<double><<<std::min<unsigned>(nblk(n, 256), 8192), 256, 0, s>>>(ypad_.as<double>(), y, n);
        LAUNCH_CHECK();
        return;
    }
    prof.begin(KernelProfile::SPMV, stream);
    k_spmv<<<nblk(nb, 256), 256, 0, s>>>(spmv_ptr_.as<int64_t>(), spmv_ent_.as<int64_t>(),
                                        spmv_col_.as<int32_t>(), Ax, x, y, nb);
    LAUNCH_CHECK();
    prof.end(stream);
}

void FactorizerDev::solve_refined(const double* Ax, const double* b, double* x, double tol,
                                  int max_iters, int* iters, double* relres, void* stream) {
    require(2);
    hipStream_t s = S_(stream);
    const int64_t nf = sym->n_full;
    scratch_c_.alloc((size_t)(3 * nf + 2 * NRED + 8) * sizeof(double));
    double* y = scratch_c_.as<double>();
    double* r = y + nf;
    double* dx = r + nf;
    double* part = dx + nf;
    double* nrm = part + NRED;   // [0] = |b|^2, [1] = |r|^2
    const int64_t* mask = f2r_.as<int64_t>();
    auto sumsq = [&](const double* v, double* out) {
        k_sumsq_partial<<<NRED, 256, 0, s>>>(v, mask, nf, part);
        LAUNCH_CHECK();
        k_sum_final<<<1, 256, 0, s>>>(part, out);
        LAUNCH_CHECK();
    };
    sumsq(b, nrm);
    solve(b, x, 1, stream);
    int it = 0;
    double h[2] = {0, 0};
    double rel = 0;
    while (true) {
        spmv(Ax, x, y, stream);
        prof.begin(KernelProfile::REFINE, s);
        k_residual<<<nblk(nf, 256), 256, 0, s>>>(b, y, mask, r, nf);
        LAUNCH_CHECK();
        sumsq(r, nrm + 1);
        prof.end(s);
        HIP_CHECK(hipMemcpyAsync(h, nrm, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        rel = h[0] > 0 ? std::sqrt(h[1] / h[0]) : std::sqrt(h[1]);
        if (rel <= tol || it >= max_iters) break;
        solve(r, dx, 1, stream);
        prof.begin(KernelProfile::REFINE, s);
        k_axpy1<<<nblk(nf, 256), 256, 0, s>>>(x, dx, nf);
        LAUNCH_CHECK();
        prof.end(s);
        ++it;
    }
    if (iters) *iters = it;
    if (relres) *relres = rel;
    if (prof.on) {
        HIP_CHECK(hipStreamSynchronize(s));
        prof.collect();
    }
}

void FactorizerDev::solve_refined_segments(const double* Ax, const double* b, double* x, int64_t nseg,
                                           const int64_t* seg, double tol, int max_iters, int* iters,
                                           double* relres_max, double* relres_seg, void* stream) {
    require(2);
    const int64_t nf = sym->n_full;
    if (nseg < 1 || nseg > (int64_t)1 << 30) throw Error(-1, "bad segment count");
    if (seg[0] != 0 || seg[nseg] != nf) throw Error(-1, "segments must cover the full vector");
    for (int64_t q = 0; q < nseg; ++q)
        if (seg[q + 1] < seg[q]) throw Error(-1, "segment offsets must be non-decreasing");
    hipStream_t s = S_(stream);
    std::vector<int64_t> sv(seg, seg + nseg + 1);
    segp_.upload(sv);
    segsum_.alloc((size_t)2 * nseg * sizeof(double));
    double* bn = segsum_.as<double>();
    double* rn = bn + nseg;
    scratch_c_.alloc((size_t)3 * nf * sizeof(double));
    double* y = scratch_c_.as<double>();
    double* r = y + nf;
    double* dx = r + nf;
    const int64_t* mask = f2r_.as<int64_t>();
    k_seg_sumsq<<<(unsigned)nseg, 256, 0, s>>>(b, mask, segp_.as<int64_t>(), bn);
    LAUNCH_CHECK();
    std::vector<double> hb(nseg), hr(nseg), rel(nseg);
    HIP_CHECK(hipMemcpyAsync(hb.data(), bn, nseg * sizeof(double), hipMemcpyDeviceToHost, s));
    solve(b, x, 1, stream);
    int it = 0;
    double worst = 0;
    while (true) {
        spmv(Ax, x, y, stream);
        prof.begin(KernelProfile::REFINE, s);
        k_residual<<<nblk(nf, 256), 256, 0, s>>>(b, y, mask, r, nf);
        LAUNCH_CHECK();
        k_seg_sumsq<<<(unsigned)nseg, 256, 0, s>>>(r, mask, segp_.as<int64_t>(), rn);
        LAUNCH_CHECK();
        prof.end(s);
        HIP_CHECK(hipMemcpyAsync(hr.data(), rn, nseg * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        worst = 0;
        for (int64_t q = 0; q < nseg; ++q) {
            rel[q] = hb[q] > 0 ? std::sqrt(hr[q] / hb[q]) : std::sqrt(hr[q]);
            worst = std::max(worst, rel[q]);
        }
        if (worst <= tol || it >= max_iters) break;
        solve(r, dx, 1, stream);
        prof.begin(KernelProfile::REFINE, s);
        k_axpy1<<<nblk(nf, 256), 256, 0, s>>>(x, dx, nf);
        LAUNCH_CHECK();
        prof.end(s);
        ++it;
    }
    if (iters) *iters = it;
    if (relres_max) *relres_max = worst;
    if (relres_seg) std::copy(rel.begin(), rel.end(), relres_seg);
    if (prof.on) {
        HIP_CHECK(hipStreamSynchronize(s));
        prof.collect();
    }
}

void FactorizerDev::autotune(const int64_t* Ap, const int64_t* Ai, int64_t nb, int bs, const int64_t* pins,
                             int64_t npins, const double* Ax, const int* fuse, int nfuse, const int* tiles, int ntile,
                             int reps, double* ms, int* best, void* stream) {
    if (bs != 3) throw Error(-6, "only uniform 3x3 blocks are supported");
    if (nfuse < 1 || ntile < 1 || reps < 1) throw Error(-1, "autotune needs at least one candidate and one repetition");
    SymbolicOptions base = options_from_env();
    base.schur_tile = schur_tile();
    std::vector<int32_t> perm;
    SepForest forest;
    {
        SymbolicOptions o0 = base;
        o0.nd_state = ordering_state(pins, npins);
        Symbolic S0 = analyze(Ap, Ai, nb, pins, npins, o0);     // the ordering is shared by all candidates
        perm = nd_order(S0);
        forest = std::move(S0.nd_forest);
    }
    base.given_forest = &forest;
    base.given_perm = &perm;
    hipStream_t s = S_(stream);
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    // fastest of `reps` numeric factorizations after one warm-up, under the schedule of `o`
    auto time_cfg = [&](const SymbolicOptions& o) {
        install(std::make_unique<Symbolic>(analyze(Ap, Ai, nb, pins, npins, o)));
        numeric(Ax, 0.0, nullptr, stream);                       // warm-up (and pivot check)
        double t = 1e300;
        for (int r = 0; r < reps; ++r) {
            HIP_CHECK(hipEventRecord(e0, s));
            numeric(Ax, 0.0, nullptr, stream);
            HIP_CHECK(hipEventRecord(e1, s));
            HIP_CHECK(hipEventSynchronize(e1));
            float m = 0;
            HIP_CHECK(hipEventElapsedTime(&m, e0, e1));
            t = std::min(t, (double)m);
        }
        return t;
    };
    // coordinate descent from the first candidates: the fused-kernel bound first, then the
    // largest GEMM tile of each size bucket; a candidate replaces the incumbent only when it is
    // at least 2 % faster, so timing noise between near-equal schedules does not flip the choice
    constexpr double kGain = 0.98;
    SymbolicOptions cur = base;
    cur.fuse_degree = std::max(0, fuse[0]);
    for (int b = 0; b < 3; ++b) cur.bucket_tile[b] = tiles[0] > 64 ? 128 : 64;
    double t_cur = time_cfg(cur);
    if (ms) ms[0] = t_cur;
    for (int i = 1; i < nfuse; ++i) {
        SymbolicOptions o = cur;
        o.fuse_degree = std::max(0, fuse[i]);
        const double t = time_cfg(o);
        if (ms) ms[i] = t;
        if (t < kGain * t_cur) { cur = o; t_cur = t; }
    }
    for (int b = 0; b < 3; ++b)
        for (int j = 0; j < ntile; ++j) {
            SymbolicOptions o = cur;
            o.bucket_tile[b] = tiles[j] > 64 ? 128 : 64;
            double t = t_cur;
            if (o.bucket_tile[b] != cur.bucket_tile[b]) {
                t = time_cfg(o);
                if (t < kGain * t_cur) { cur = o; t_cur = t; }
            }
            if (ms) ms[nfuse + b * ntile + j] = t;
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    install(std::make_unique<Symbolic>(analyze(Ap, Ai, nb, pins, npins, cur)));
    if (best) {
        best[0] = cur.fuse_degree;
        for (int b = 0; b < 3; ++b) best[1 + b] = cur.bucket_tile[b];
    }
}

// L of P A P^T (for an FP32 factor: the factor of S P A P^T S, unscaled to L = S^{-1} L~)
void FactorizerDev::export_csc(int64_t* Lp, int64_t* Li, double* Lx) {
    require(2);
    const Symbolic& S = *sym;
    std::vector<double> h(S.L_size);
    HIP_CHECK(hipDeviceSynchronize());
    if (precision != 0) {
        std::vector<float> hf(S.L_size);
        HIP_CHECK(hipMemcpy(hf.data(), L_.p, S.L_size * sizeof(float), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < S.L_size; ++i) h[i] = hf[i];
    } else {
        HIP_CHECK(hipMemcpy(h.data(), L_.p, S.L_size * sizeof(double), hipMemcpyDeviceToHost));
    }
    panels_to_csc(S, h, Lp, Li, Lx);
    if (precision != 0) {
        std::vector<double> sc(3 * S.n_red);
        HIP_CHECK(hipMemcpy(sc.data(), scale_.p, sc.size() * sizeof(double), hipMemcpyDeviceToHost));
        const int64_t n = S.n_free;
        for (int64_t j = 0; j < n; ++j)
            for (int64_t q = Lp[j]; q < Lp[j + 1]; ++q) Lx[q] /= sc[S.uperm_to_int[Li[q]]];
    }
}

void FactorizerDev::clear_factors() {
    if (ftype > 0) ftype = 0;
}

void FactorizerDev::stash() {
    require(1);
    Lstash_.alloc(L_.bytes);
    HIP_CHECK(hipMemcpy(Lstash_.p, L_.p, L_.bytes, hipMemcpyDeviceToDevice));
    invstash_.alloc(inv_.bytes);
    HIP_CHECK(hipMemcpy(invstash_.p, inv_.p, inv_.bytes, hipMemcpyDeviceToDevice));
    xbstash_.alloc(Xb_.bytes);
    HIP_CHECK(hipMemcpy(xbstash_.p, Xb_.p, Xb_.bytes, hipMemcpyDeviceToDevice));
    if (precision != 0) {
        scalestash_.alloc(scale_.bytes);
        HIP_CHECK(hipMemcpy(scalestash_.p, scale_.p, scale_.bytes, hipMemcpyDeviceToDevice));
    }
    stash_valid_ = true;
    stash_type_ = ftype;
    stash_succ_ = num_successful;
}

void FactorizerDev::swap_stash() {
    if (!stash_valid_) throw Error(-2, "No stashed factorization");
    drop_graph();                 // the captured graph addresses the current panels
    // the panels, the tile and block inverses and (FP32) the scaling all belong to the factor
    std::swap(L_.p, Lstash_.p);
    std::swap(L_.bytes, Lstash_.bytes);
    std::swap(inv_.p, invstash_.p);
    std::swap(inv_.bytes, invstash_.bytes);
    std::swap(Xb_.p, xbstash_.p);
    std::swap(Xb_.bytes, xbstash_.bytes);
    if (precision != 0) {
        std::swap(scale_.p, scalestash_.p);
        std::swap(scale_.bytes, scalestash_.bytes);
    }
    std::swap(ftype, stash_type_);
    std::swap(num_successful, stash_succ_);
    if (ftype == 0) ftype = 1;   // the symbolic analysis is shared by both factors
}

void FactorizerDev::clear_stash() {
    Lstash_.release();
    invstash_.release();
    xbstash_.release();
    scalestash_.release();
    stash_valid_ = false;
}

// ------------------------------------------------------------------------------------------
// test hooks
// ------------------------------------------------------------------------------------------
template <typename T>
void test_gemm_nt(int M, int N, int K, T alpha, const T* A, int lda, const T* B, int ldb, T beta, T* C,
                  int ldc, void* stream) {
    if (!(alpha == T(1) || alpha == T(-1)) || !(beta == T(0) || beta == T(1)))
        throw Error(-1, "the tile kernel supports alpha in {+1,-1} and beta in {0,1}");
    if (M < 0 || N < 0 || K < 0 || lda < std::max(1, M) || ldb < std::max(1, N) || ldc < std::max(1, M))
        throw Error(-1, "bad GEMM dimensions");
    // 128x128 tiles when both dimensions reach 128 (as the factorization schedule does), else 64x64
    const int TT = (M >= 128 && N >= 128) ? 128 : 64;
    std::vector<GemmTask> tasks;
    for (int c0 = 0; c0 < N; c0 += TT)
        for (int r0 = 0; r0 < M; r0 += TT) {
            GemmTask g{};
            g.a_off = r0;
            g.lda = lda;
            g.b_off = c0;
            g.ldb = ldb;
            g.c_off = (int64_t)c0 * ldc + r0;
            g.ldc = ldc;
            g.M = std::min(TT, M - r0);
            g.N = std::min(TT, N - c0);
            g.K = K;
            g.flags = GT_C_IN_U | GT_B_IN_INV | (beta == T(0) ? GT_BETA0 : 0) | (alpha > T(0) ? GT_ALPHA_POS : 0);
            tasks.push_back(g);
        }
    if (tasks.empty()) return;
    DeviceArray d;
    d.upload(tasks);
    hipStream_t s = S_(stream);
    TaskRange r;
    r.b = 0;
    r.m = TT == 128 ? (int64_t)tasks.size() : 0;
    r.e = (int64_t)tasks.size();
    // the operands in arrays with the factor arrays' tails (the direct-to-LDS kernel reads up to a
    // tile height past a column's start)
    DeviceArray ap, bp;
    const size_t ab = (size_t)std::max(K, 1) * lda * sizeof(T), bb = (size_t)std::max(K, 1) * ldb * sizeof(T);
    ap.alloc(ab + 256 * sizeof(T));
    bp.alloc(bb + 256 * sizeof(T));
    HIP_CHECK(hipMemsetAsync(ap.p, 0, ap.bytes, s));
    HIP_CHECK(hipMemsetAsync(bp.p, 0, bp.bytes, s));
    if (K > 0) {
        HIP_CHECK(hipMemcpyAsync(ap.p, A, (size_t)K * lda * sizeof(T), hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync(bp.p, B, (size_t)K * ldb * sizeof(T), hipMemcpyDeviceToDevice, s));
    }
    launch_gemm<T>(r, d.as<GemmTask>(), ap.as<T>(), C, bp.as<T>(), s);
    HIP_CHECK(hipStreamSynchronize(s));
}
void test_gemm_nt_bf16(int M, int N, int K, float alpha, const float* A, int lda, const float* B, int ldb,
                       float beta, float* C, int ldc, bool split, void* stream) {
    if (!(alpha == 1.0f || alpha == -1.0f) || !(beta == 0.0f || beta == 1.0f))
        throw Error(-1, "the tile kernel supports alpha in {+1,-1} and beta in {0,1}");
    if (M < 0 || N < 0 || K < 0 || lda < std::max(1, M) || ldb < std::max(1, N) || ldc < std::max(1, M))
        throw Error(-1, "bad GEMM dimensions");
    const int TT = (M >= 256 && N >= 256) ? 256 : (M >= 128 && N >= 128) ? 128 : 64;
    std::vector<GemmTask> tasks;
    for (int c0 = 0; c0 < N; c0 += TT)
        for (int r0 = 0; r0 < M; r0 += TT) {
            GemmTask g{};
            g.a_off = r0;
            g.lda = lda;
            g.b_off = c0;
            g.ldb = ldb;
            g.c_off = (int64_t)c0 * ldc + r0;
            g.ldc = ldc;
            g.M = std::min(TT, M - r0);
            g.N = std::min(TT, N - c0);
            g.K = K;
            g.flags = GT_C_IN_U | GT_B_IN_INV | (beta == 0.0f ? GT_BETA0 : 0) | (alpha > 0.0f ? GT_ALPHA_POS : 0);
            tasks.push_back(g);
        }
    if (tasks.empty()) return;
    DeviceArray d;
    d.upload(tasks);
    hipStream_t s = S_(stream);
    TaskRange r;
    r.b = 0;
    r.m = TT >= 128 ? (int64_t)tasks.size() : 0;
    r.e = (int64_t)tasks.size();
    launch_gemm_bf(r, d.as<GemmTask>(), const_cast<float*>(A), C, B, split, TT == 256, s);
    HIP_CHECK(hipStreamSynchronize(s));
}

template void test_gemm_nt<double>(int, int, int, double, const double*, int, const double*, int, double,
                                   double*, int, void*);
template void test_gemm_nt<float>(int, int, int, float, const float*, int, const float*, int, float,
                                  float*, int, void*);

void test_multiply_lower_block(int BS, const int64_t* indices, int64_t start, int64_t size,
                               int64_t degree, const double* A, int64_t lda, int nrhs, double* B,
                               int64_t ldb, void* stream) {
    if (BS < 1 || BS > 3 || size < 0 || degree < 0 || nrhs < 1 || degree % BS)
        throw Error(-1, "bad MultiplyLowerBlock arguments");
    if (size == 0 || degree == 0) return;
    hipStream_t s = S_(stream);
    dim3 g(1, (unsigned)nrhs);
    k_mlb<<<g, 256, (size_t)size * sizeof(double), s>>>(BS, indices, start, size, degree, A, lda, B, ldb);
    LAUNCH_CHECK();
    HIP_CHECK(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------------------------------
// Allocation-free solve: its scratch is this file-scope device array (one copy per GPU, held
// until the process ends).  Host code never passes the array's name to a kernel: that is a
// host-side stand-in, not a device address.  The device address is looked up once per device
// with hipGetSymbolAddress and only that pointer is used.  Every allocation-free solve on one
// GPU shares the array, so calls that overlap on one GPU race.
// ------------------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[MFS_SOLVE_NOALLOC_MAX_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}

void* noalloc_scratch() {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(-4, "device index out of range for the solve scratch");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(-4, std::string("hipGetSymbolAddress of the solve scratch failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

} // namespace mfs
