// On-disk format of BlockCSCHessian (BlockCSCHessian.cc:11-103): size_t blockSize, size_t
// numBlocks, then int64 m, n (block rows/columns), nz (stored blocks), snz (scalar values),
// uint32 symmetry mode (1 = upper triangle), Ap[n + 1], Ai[nz] (int64), Ax[snz] (double).  The file
// always holds the non-contiguous layout, whose Ax is the scalar CSC of the upper triangle: scalar
// column 3 bj + c lists the three rows of every off-diagonal block of block column bj, then rows
// 0..c of its diagonal block.  The device path uses the contiguous layout (9 values per block), so
// reading and writing convert.
#include "mfs_internal.h"

#include <cstdio>
#include <memory>

namespace mfs {

namespace {
struct File {
    FILE* f;
    explicit File(const char* p, const char* mode) : f(fopen(p, mode)) {
        if (!f) throw Error(-1, std::string("cannot open ") + p);
    }
    ~File() { if (f) fclose(f); }
    void put(const void* d, size_t b) {
        if (b && fwrite(d, 1, b, f) != b) throw Error(-1, "write failed");
    }
    void get(void* d, size_t b) {
        if (b && fread(d, 1, b, f) != b) throw Error(-1, "truncated block-CSC file");
    }
};

void check_pattern(const int64_t* Ap, const int64_t* Ai, int64_t nb) {
    if (Ap[0] != 0) throw Error(-1, "Ap[0] must be 0");
    for (int64_t j = 0; j < nb; ++j) {
        if (Ap[j + 1] <= Ap[j]) throw Error(-1, "every block column needs its diagonal block");
        for (int64_t k = Ap[j]; k < Ap[j + 1]; ++k) {
            if (Ai[k] < 0 || Ai[k] > j || (k > Ap[j] && Ai[k] <= Ai[k - 1]))
                throw Error(-1, "pattern must be sorted upper-triangular block CSC");
        }
        if (Ai[Ap[j + 1] - 1] != j) throw Error(-1, "diagonal block must be last in its column");
    }
}
} // namespace

void write_blockcsc(const char* path, const int64_t* Ap, const int64_t* Ai, int64_t nb, const double* Ax) {
    check_pattern(Ap, Ai, nb);
    const int64_t nz = Ap[nb];
    const int64_t snz = Ax ? 9 * nz - 3 * nb : 0;
    File F(path, "wb");
    const size_t bs = 3, nblk = (size_t)nb;
    F.put(&bs, sizeof bs);
    F.put(&nblk, sizeof nblk);
    const int64_t hdr[4] = {nb, nb, nz, snz};
    F.put(hdr, sizeof hdr);
    const uint32_t sym = 1;
    F.put(&sym, sizeof sym);
    F.put(Ap, (nb + 1) * sizeof(int64_t));
    F.put(Ai, nz * sizeof(int64_t));
    if (!Ax) return;
    std::vector<double> col;
    for (int64_t j = 0; j < nb; ++j) {
        const int64_t k0 = Ap[j], kd = Ap[j + 1] - 1;
        for (int c = 0; c < 3; ++c) {
            col.clear();
            for (int64_t k = k0; k < kd; ++k)
                for (int r = 0; r < 3; ++r) col.push_back(Ax[9 * k + 3 * c + r]);
            for (int r = 0; r <= c; ++r) col.push_back(Ax[9 * kd + 3 * c + r]);
            F.put(col.data(), col.size() * sizeof(double));
        }
    }
}

// Reads a file written in the format above.  With null outputs only the sizes are returned.
void read_blockcsc(const char* path, int64_t* nb_out, int64_t* nz_out, int64_t* snz_out, int64_t* Ap, int64_t* Ai,
                   double* Ax) {
    File F(path, "rb");
    size_t bs = 0, nblk = 0;
    F.get(&bs, sizeof bs);
    F.get(&nblk, sizeof nblk);
    if (bs != 3) throw Error(-6, "only block size 3 is supported");
    int64_t hdr[4];
    F.get(hdr, sizeof hdr);
    uint32_t sym = 0;
    F.get(&sym, sizeof sym);
    if (sym > 2) throw Error(-1, "Invalid symmetry_mode: " + std::to_string(sym));
    const int64_t m = hdr[0], n = hdr[1], nz = hdr[2], snz = hdr[3];
    if (m != n || n != (int64_t)nblk || n <= 0 || nz < n) throw Error(-1, "inconsistent block-CSC header");
    if (snz != 0 && snz != 9 * nz - 3 * n) throw Error(-1, "constructFromBinaryStream: scalarNNZ mismatch");
    if (nb_out) *nb_out = n;
    if (nz_out) *nz_out = nz;
    if (snz_out) *snz_out = snz;
    if (!Ap && !Ai && !Ax) return;
    std::vector<int64_t> p(n + 1), i(nz);
    F.get(p.data(), p.size() * sizeof(int64_t));
    F.get(i.data(), i.size() * sizeof(int64_t));
    if (p[n] != nz) throw Error(-1, "inconsistent block-CSC header");
    check_pattern(p.data(), i.data(), n);
    if (Ap) std::copy(p.begin(), p.end(), Ap);
    if (Ai) std::copy(i.begin(), i.end(), Ai);
    if (!Ax) return;
    if (snz == 0) throw Error(-1, "the file holds a sparsity pattern only");
    std::vector<double> col;
    for (int64_t j = 0; j < n; ++j) {
        const int64_t k0 = p[j], kd = p[j + 1] - 1;
        for (int c = 0; c < 3; ++c) {
            col.resize(3 * (kd - k0) + c + 1);
            F.get(col.data(), col.size() * sizeof(double));
            size_t q = 0;
            for (int64_t k = k0; k < kd; ++k)
                for (int r = 0; r < 3; ++r) Ax[9 * k + 3 * c + r] = col[q++];
            for (int r = 0; r <= c; ++r) Ax[9 * kd + 3 * c + r] = col[q++];
            for (int r = c + 1; r < 3; ++r) Ax[9 * kd + 3 * c + r] = 0.0;   // strict lower of the diagonal block
        }
    }
}

} // namespace mfs
