// Runs the library's MINRES-QLP (csrc/minresqlp.hpp: the launch/finish form the SR
// step uses, fused update-and-reduction passes, rotated buffers) on a dense Hermitian
// system read from a file, with the operator computed in the oracle's order
// (out[r] = sum_q A[r][q] v[q], q ascending), and writes x, the flag and the iteration
// count.  Input file: n (int64), rtol, maxit (double, uint64), A (n x n complex128,
// row-major), b (n complex128).  Output file: flag (int64), iters (uint64), x.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "minresqlp.hpp"

using nq64::cd;

int main(int argc, char **argv)
{
    if (argc != 3) return 2;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    int64_t n = 0;
    double rtol = 0;
    uint64_t maxit = 0;
    if (std::fread(&n, 8, 1, f) != 1 || std::fread(&rtol, 8, 1, f) != 1 || std::fread(&maxit, 8, 1, f) != 1) return 4;
    std::vector<cd> A((size_t)(n * n)), b((size_t)n), x;
    if (std::fread(A.data(), sizeof(cd), A.size(), f) != A.size() || std::fread(b.data(), sizeof(cd), b.size(), f) != b.size())
        return 4;
    std::fclose(f);
    auto op = [&](const std::vector<cd> &v, std::vector<cd> &out) -> bool {
        for (int64_t r = 0; r < n; r++) {
            cd acc = nq64::zero();
            for (int64_t q = 0; q < n; q++) acc = nq64::add(acc, nq64::mul(A[(size_t)(r * n + q)], v[(size_t)q]));
            out[(size_t)r] = acc;
        }
        return true;
    };
    bool aborted = false;
    const nqmq::Result res = nqmq::solve(op, b, x, rtol, maxit, aborted);
    if (aborted) return 5;
    FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 3;
    const int64_t flag = res.flag;
    std::fwrite(&flag, 8, 1, o);
    std::fwrite(&res.iters, 8, 1, o);
    std::fwrite(x.data(), sizeof(cd), x.size(), o);
    std::fclose(o);
    return 0;
}
