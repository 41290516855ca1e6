// Test driver: reads a MAT v5 file with the product's reader, prints one line per
// variable and writes all variables back to a second file with the product's writer.
#include <cstdio>

#include "matv5.hpp"

int main(int argc, char **argv)
{
    if (argc != 3) return 2;
    try {
        auto vars = matv5::read_file(argv[1]);
        for (auto &v : vars) {
            printf("%s %d %d", v.name.c_str(), v.complex_ ? 1 : 0, (int)v.dims.size());
            for (auto d : v.dims) printf(" %lld", (long long)d);
            printf("\n");
        }
        matv5::write_file(argv[2], vars);
    } catch (const std::exception &e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
